/*
 * lightmetrica_mirror.h -- the reference's component interfaces on the hot path, mirrored
 * for hosts and plugins built without Boost.
 *
 * Names, virtual signatures and semantics follow the reference headers (paths relative to
 * the reference root):
 *   Component / ComponentFactory ........ include/lightmetrica/component.h:45-197
 *   plugin entry points .................. src/liblightmetrica/plugin.common.cpp:97-109,
 *                                          plugin file names ^plugin\.([a-z\.]+)\.so$ (component.cpp:69)
 *   GeneralizedBSDF queries / types ...... include/lightmetrica/generalizedbsdf.h:34-246
 *   SurfaceGeometry ...................... include/lightmetrica/surfacegeometry.h:40-62
 *   Emitter / Light / Camera / BSDF ...... emitter.h:43-98, light.h:42-51, camera.h:45-66, bsdf.h:33-85
 *   TriangleMesh / Primitive / Primitives  trianglemesh.h:44-82, primitive.h:39-60, primitives.h:52-137
 *   Film / BitmapFilm / BitmapImage ...... film.h:34-100, bitmapfilm.h:45-99, bitmap.h:34-84
 *   Scene ................................ include/lightmetrica/scene.h:47-191, src/liblightmetrica/scene.cpp
 *   Renderer / RenderProcess ............. renderer.h:39-133, renderproc.h:41-96
 *   RenderProcessScheduler ............... sched.h:52-125
 *
 * Differences: the Connect_Report* members that return boost::signals2 connections are
 * replaced by SetReportProgress(std::function); ConfigNode is a small in-memory tree
 * instead of the pugixml-backed one; Math::Float is float (the reference's default
 * single-precision build, CMakeLists.txt:155).
 */
#ifndef LIGHTMETRICA_AMD_MIRROR_H
#define LIGHTMETRICA_AMD_MIRROR_H

#include <dirent.h>
#include <dlfcn.h>

#include <cmath>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <regex>
#include <sstream>
#include <string>
#include <vector>

namespace lightmetrica {

// ---------------------------------------------------------------------------------------
// Math (single precision)

namespace Math {

using Float = float;

struct Vec2 {
    Float x = 0, y = 0;
    Vec2() {}
    explicit Vec2(Float s) : x(s), y(s) {}
    Vec2(Float a, Float b) : x(a), y(b) {}
    Float operator[](int i) const { return i == 0 ? x : y; }
};

struct Vec3 {
    Float x = 0, y = 0, z = 0;
    Vec3() {}
    explicit Vec3(Float s) : x(s), y(s), z(s) {}
    Vec3(Float a, Float b, Float c) : x(a), y(b), z(c) {}
    Float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    Float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    Vec3 operator+(const Vec3& o) const { return Vec3(x + o.x, y + o.y, z + o.z); }
    Vec3 operator-(const Vec3& o) const { return Vec3(x - o.x, y - o.y, z - o.z); }
    Vec3 operator-() const { return Vec3(-x, -y, -z); }
    Vec3 operator*(Float s) const { return Vec3(x * s, y * s, z * s); }
    Vec3 operator*(const Vec3& o) const { return Vec3(x * o.x, y * o.y, z * o.z); }
    Vec3 operator/(Float s) const { return Vec3(x / s, y / s, z / s); }
};

struct Vec4 {
    Float x = 0, y = 0, z = 0, w = 0;
    Vec4() {}
    Vec4(Float a, Float b, Float c, Float d) : x(a), y(b), z(c), w(d) {}
    Vec4(const Vec3& v, Float d) : x(v.x), y(v.y), z(v.z), w(d) {}
    Float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : (i == 2 ? z : w)); }
    Float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : (i == 2 ? z : w)); }
};

// Column-major: v[i] is column i (math.matrix.inl:343-350).
struct Mat3 {
    Vec3 v[3];
    Mat3() {}
    Mat3(const Vec3& a, const Vec3& b, const Vec3& c) { v[0] = a; v[1] = b; v[2] = c; }
    Vec3 operator*(const Vec3& p) const {
        return Vec3(v[0].x * p.x + v[1].x * p.y + v[2].x * p.z, v[0].y * p.x + v[1].y * p.y + v[2].y * p.z,
                    v[0].z * p.x + v[1].z * p.y + v[2].z * p.z);
    }
};

struct Mat4 {
    Vec4 v[4];
    static Mat4 Identity() {
        Mat4 m;
        for (int i = 0; i < 4; i++) m.v[i][i] = 1;
        return m;
    }
    Vec4 operator*(const Vec4& p) const {
        Vec4 r;
        for (int row = 0; row < 4; row++)
            r[row] = v[0][row] * p.x + v[1][row] * p.y + v[2][row] * p.z + v[3][row] * p.w;
        return r;
    }
};

inline Float Dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 Cross(const Vec3& a, const Vec3& b) {
    return Vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline Float Length(const Vec3& a) { return std::sqrt(Dot(a, a)); }
inline Vec3 Normalize(const Vec3& a) { return a / Length(a); }
inline bool IsZero(const Vec3& a) { return a.x == 0 && a.y == 0 && a.z == 0; }

// Inverse of a 4x4 matrix by Gauss-Jordan elimination in double.
inline Mat4 Inverse(const Mat4& m) {
    double a[4][8];
    for (int r = 0; r < 4; r++) {
        for (int c = 0; c < 4; c++) a[r][c] = m.v[c][r];
        for (int c = 0; c < 4; c++) a[r][4 + c] = r == c ? 1.0 : 0.0;
    }
    for (int c = 0; c < 4; c++) {
        int p = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        for (int k = 0; k < 8; k++) std::swap(a[c][k], a[p][k]);
        const double iv = 1.0 / a[c][c];
        for (int k = 0; k < 8; k++) a[c][k] *= iv;
        for (int r = 0; r < 4; r++) {
            if (r == c) continue;
            const double f = a[r][c];
            for (int k = 0; k < 8; k++) a[r][k] -= f * a[c][k];
        }
    }
    Mat4 out;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) out.v[c][r] = Float(a[r][4 + c]);
    return out;
}

// Transpose(Inverse(m)) restricted to 3x3 -- Primitive::normalTransform (primitive.h:44)
inline Mat3 NormalMatrix(const Mat4& m) {
    const Mat4 inv = Inverse(m);
    Mat3 n;
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) n.v[c][r] = inv.v[r][c];
    return n;
}

enum class ProbabilityMeasure { None, SolidAngle, ProjectedSolidAngle, Area, Discrete };
struct PDFEval {
    Float v = 0;
    ProbabilityMeasure measure = ProbabilityMeasure::None;
    PDFEval() {}
    PDFEval(Float value, ProbabilityMeasure m) : v(value), measure(m) {}
};

constexpr Float Pi() { return 3.14159265358979323846f; }
constexpr Float InvPi() { return 0.318309886183790671538f; }

}  // namespace Math

// ---------------------------------------------------------------------------------------
// Components and plugins

class Component {
  public:
    virtual ~Component() {}
    virtual std::string ComponentInterfaceTypeName() const = 0;
    virtual std::string ComponentImplTypeName() const = 0;
};

#define LM_COMPONENT_INTERFACE_DEF(Name)                       \
    static const char* InterfaceTypeName() { return Name; }   \
    virtual std::string ComponentInterfaceTypeName() const override { return InterfaceTypeName(); }

#define LM_COMPONENT_IMPL_DEF(Name)                       \
    static const char* ImplTypeName() { return Name; }   \
    virtual std::string ComponentImplTypeName() const override { return ImplTypeName(); }

using PluginCreateInstanceFunc = Component* (*)(const char* implType, const char* interfaceType);
using PluginCheckRegisteredFunc = bool (*)(const char* implType, const char* interfaceType);

/*
 * Host-side factory restricted to plugins: LoadPlugins dlopens every file of a directory
 * whose name matches ^plugin\.([a-z\.]+)\.so$ (component.cpp:61-120, dynamiclibrary.cpp:78)
 * and Create<I>(impl) asks each plugin in load order (component.cpp:152-165), returning
 * nullptr for an unknown type (component.h:168-180).
 */
class ComponentFactory {
  public:
    struct Plugin {
        std::string path;
        void* handle;
        PluginCreateInstanceFunc create;
        PluginCheckRegisteredFunc check;
    };

    static std::vector<Plugin>& Plugins() {
        static std::vector<Plugin> plugins;
        return plugins;
    }

    static int LoadPlugins(const std::string& dir, std::string* log = nullptr) {
        DIR* d = opendir(dir.c_str());
        if (!d) return 0;
        const std::regex pattern("^plugin\\.([a-z\\.]+)\\.so$");
        int loaded = 0;
        while (dirent* e = readdir(d)) {
            const std::string name = e->d_name;
            if (!std::regex_match(name, pattern)) continue;
            const std::string path = dir + "/" + name;
            void* h = dlopen(path.c_str(), RTLD_LAZY | RTLD_LOCAL);
            if (!h) {
                if (log) *log += std::string("failed to load ") + path + ": " + dlerror() + "\n";
                continue;
            }
            auto create = reinterpret_cast<PluginCreateInstanceFunc>(dlsym(h, "LM_Plugin_CreateInstance"));
            auto check = reinterpret_cast<PluginCheckRegisteredFunc>(dlsym(h, "LM_Plugin_CheckRegistered"));
            if (!create || !check) {
                if (log) *log += "missing plugin entry points in " + path + "\n";
                dlclose(h);
                continue;
            }
            Plugins().push_back(Plugin{path, h, create, check});
            loaded++;
        }
        closedir(d);
        return loaded;
    }

    static void UnloadPlugins() {
        for (auto& p : Plugins()) dlclose(p.handle);
        Plugins().clear();
    }

    static bool CheckRegistered(const std::string& interfaceType, const std::string& implType) {
        for (auto& p : Plugins())
            if (p.check(implType.c_str(), interfaceType.c_str())) return true;
        return false;
    }

    static Component* Create(const std::string& interfaceType, const std::string& implType) {
        for (auto& p : Plugins())
            if (p.check(implType.c_str(), interfaceType.c_str())) return p.create(implType.c_str(), interfaceType.c_str());
        return nullptr;
    }

    template <class InterfaceType> static InterfaceType* Create(const std::string& implType) {
        Component* c = Create(InterfaceType::InterfaceTypeName(), implType);
        InterfaceType* p = dynamic_cast<InterfaceType*>(c);
        if (c && !p) delete c;
        return p;
    }
};

// ---------------------------------------------------------------------------------------
// Configuration: a small tree of named text values (ConfigNode, confignode.h:139-160).

class ConfigNode {
  public:
    ConfigNode() {}
    ConfigNode(const std::string& name, const std::string& value = "") : name_(name), value_(value) {}
    ConfigNode& Add(const std::string& name, const std::string& value) {
        children_.push_back(std::make_shared<ConfigNode>(name, value));
        return *children_.back();
    }
    bool Empty() const { return name_.empty(); }
    const std::string& Name() const { return name_; }
    const std::string& Value() const { return value_; }
    ConfigNode Child(const std::string& name) const {
        for (auto& c : children_)
            if (c->name_ == name) return *c;
        return ConfigNode();
    }
    template <class T> bool ChildValue(const std::string& name, T& out) const {
        ConfigNode c = Child(name);
        if (c.Empty()) return false;
        std::istringstream ss(c.value_);
        T v;
        if (!(ss >> v)) return false;
        out = v;
        return true;
    }
    template <class T> bool ChildValueOrDefault(const std::string& name, const T& def, T& out) const {
        if (ChildValue(name, out)) return true;
        out = def;
        return false;
    }

  private:
    std::string name_, value_;
    std::vector<std::shared_ptr<ConfigNode>> children_;
};

template <> inline bool ConfigNode::ChildValue<std::string>(const std::string& name, std::string& out) const {
    ConfigNode c = Child(name);
    if (c.Empty()) return false;
    out = c.value_;
    return true;
}

class Assets {};

// ---------------------------------------------------------------------------------------
// Surface geometry, BSDF queries

enum class TransportDirection { EL = 0, LE = 1 };

enum GeneralizedBSDFType {
    DiffuseReflection = 1 << 0,
    DiffuseTransmission = 1 << 1,
    SpecularReflection = 1 << 2,
    SpecularTransmission = 1 << 3,
    GlossyReflection = 1 << 4,
    GlossyTransmission = 1 << 5,
    DeltaLightDirection = 1 << 6,
    NonDeltaLightDirection = 1 << 7,
    DeltaEyeDirection = 1 << 8,
    NonDeltaEyeDirection = 1 << 9,
    Diffuse = DiffuseReflection | DiffuseTransmission,
    Specular = SpecularReflection | SpecularTransmission,
    Glossy = GlossyReflection | GlossyTransmission,
    LightDirection = DeltaLightDirection | NonDeltaLightDirection,
    EyeDirection = DeltaEyeDirection | NonDeltaEyeDirection,
    AllEmitter = LightDirection | EyeDirection,
    AllBSDF = Diffuse | Specular | Glossy,
    All = AllEmitter | AllBSDF,
    Delta = Specular | DeltaLightDirection | DeltaEyeDirection,
    NonDelta = All & ~Delta
};

struct SurfaceGeometry {
    bool degenerated = false;
    Math::Vec3 p, gn, sn, ss, st;
    Math::Vec2 uv;
    Math::Mat3 worldToShading, shadingToWorld;
    // OrthonormalBasis(sn, ss, st) (math.linalgebra.inl:26-33), surfacegeometry.h:54-62
    void ComputeTangentSpace() {
        st = std::fabs(sn.x) > std::fabs(sn.y) ? Math::Normalize(Math::Vec3(sn.z, 0, -sn.x))
                                               : Math::Normalize(Math::Vec3(0, sn.z, -sn.y));
        ss = Math::Normalize(Math::Cross(st, sn));
        shadingToWorld = Math::Mat3(ss, st, sn);
        worldToShading = Math::Mat3(Math::Vec3(ss.x, st.x, sn.x), Math::Vec3(ss.y, st.y, sn.y),
                                    Math::Vec3(ss.z, st.z, sn.z));
    }
};

struct GeneralizedBSDFSampleQuery {
    Math::Vec2 sample;
    Math::Float uComp = 0;
    int type = 0;
    TransportDirection transportDir = TransportDirection::EL;
    Math::Vec3 wi;
};

struct GeneralizedBSDFSampleResult {
    int sampledType = 0;
    Math::Vec3 wo;
    Math::PDFEval pdf;
};

struct GeneralizedBSDFEvaluateQuery {
    int type = 0;
    TransportDirection transportDir = TransportDirection::EL;
    Math::Vec3 wi, wo;
    bool forced = false;
    GeneralizedBSDFEvaluateQuery() {}
    GeneralizedBSDFEvaluateQuery(int t, TransportDirection d, const Math::Vec3& i, const Math::Vec3& o)
        : type(t), transportDir(d), wi(i), wo(o) {}
};

// ---------------------------------------------------------------------------------------
// Assets

class Asset : public Component {
  public:
    virtual ~Asset() {}
    const std::string& ID() const { return id_; }
    void SetID(const std::string& id) { id_ = id; }

  private:
    std::string id_;
};

class GeneralizedBSDF : public Asset {
  public:
    virtual bool SampleDirection(const GeneralizedBSDFSampleQuery& query, const SurfaceGeometry& geom,
                                 GeneralizedBSDFSampleResult& result) const = 0;
    virtual Math::Vec3 SampleAndEstimateDirection(const GeneralizedBSDFSampleQuery& query, const SurfaceGeometry& geom,
                                                  GeneralizedBSDFSampleResult& result) const = 0;
    virtual Math::Vec3 EvaluateDirection(const GeneralizedBSDFEvaluateQuery& query,
                                         const SurfaceGeometry& geom) const = 0;
    virtual Math::PDFEval EvaluateDirectionPDF(const GeneralizedBSDFEvaluateQuery& query,
                                               const SurfaceGeometry& geom) const = 0;
    virtual int BSDFTypes() const = 0;
};

class BSDF : public GeneralizedBSDF {
  public:
    LM_COMPONENT_INTERFACE_DEF("bsdf");
};

class Emitter : public GeneralizedBSDF {
  public:
    virtual void SamplePosition(const Math::Vec2& sample, SurfaceGeometry& geom, Math::PDFEval& pdf) const = 0;
    virtual Math::Vec3 EvaluatePosition(const SurfaceGeometry& geom) const = 0;
    virtual Math::PDFEval EvaluatePositionPDF(const SurfaceGeometry& geom) const = 0;
};

class Light : public Emitter {
  public:
    LM_COMPONENT_INTERFACE_DEF("light");
    virtual bool EnvironmentLight() const = 0;
};

class Film;

class Camera : public Emitter {
  public:
    LM_COMPONENT_INTERFACE_DEF("camera");
    virtual bool RayToRasterPosition(const Math::Vec3& p, const Math::Vec3& d, Math::Vec2& rasterPos) const = 0;
    virtual Film* GetFilm() const = 0;
};

class TriangleMesh : public Asset {
  public:
    LM_COMPONENT_INTERFACE_DEF("trianglemesh");
    virtual int NumVertices() const = 0;
    virtual int NumFaces() const = 0;  // number of face indices (3 per triangle)
    virtual const Math::Float* Positions() const = 0;
    virtual const Math::Float* Normals() const = 0;
    virtual const Math::Float* TexCoords() const = 0;
    virtual const unsigned int* Faces() const = 0;
};

class BitmapImage {
  public:
    std::vector<Math::Float>& InternalData() { return data_; }
    const std::vector<Math::Float>& InternalData() const { return data_; }

  private:
    std::vector<Math::Float> data_;
};

class Film : public Asset {
  public:
    LM_COMPONENT_INTERFACE_DEF("film");
    virtual int Width() const = 0;
    virtual int Height() const = 0;
    virtual void RecordContribution(const Math::Vec2& rasterPos, const Math::Vec3& contrb) = 0;
    virtual void AccumulateContribution(const Math::Vec2& rasterPos, const Math::Vec3& contrb) = 0;
    virtual void AccumulateContribution(const Film& film) = 0;
    virtual void Rescale(const Math::Float& weight) = 0;
    virtual Film* Clone() const = 0;
    virtual void Clear() = 0;
};

class BitmapFilm : public Film {
  public:
    virtual void Allocate(int width, int height) = 0;
    virtual BitmapImage& Bitmap() = 0;
    virtual const BitmapImage& Bitmap() const = 0;
};

struct Primitive {
    explicit Primitive(const Math::Mat4& t) : transform(t), normalTransform(Math::NormalMatrix(t)) {}
    Math::Mat4 transform;
    Math::Mat3 normalTransform;
    TriangleMesh* mesh = nullptr;
    BSDF* bsdf = nullptr;
    Camera* camera = nullptr;
    Light* light = nullptr;
};

struct Ray {
    Math::Vec3 o, d;
    Math::Float minT = 0, maxT = 0;
};

struct AABB {
    Math::Vec3 min, max;
};

struct Intersection {
    const BSDF* bsdf = nullptr;
    const Camera* camera = nullptr;
    const Light* light = nullptr;
    SurfaceGeometry geom;
};

class Scene;

class Primitives : public Component {
  public:
    LM_COMPONENT_INTERFACE_DEF("primitives");
    virtual int NumPrimitives() const = 0;
    virtual const Primitive* PrimitiveByIndex(int index) const = 0;
    virtual const Camera* MainCamera() const = 0;
    virtual int NumLights() const = 0;
    virtual const Light* LightByIndex(int index) const = 0;
};

// ---------------------------------------------------------------------------------------
// Scene (scene.h:47-191, scene.cpp)

class Scene : public Component {
  public:
    LM_COMPONENT_INTERFACE_DEF("scene");
    virtual ~Scene() {}

    // Scene::Load takes ownership (scene.h:66-69)
    void Load(Primitives* p) { primitives_.reset(p); }
    const Camera* MainCamera() const { return primitives_ ? primitives_->MainCamera() : nullptr; }
    const Primitives* GetPrimitives() const { return primitives_.get(); }

    virtual bool Configure(const ConfigNode& node) = 0;
    virtual bool Build() = 0;
    virtual bool IntersectTriangles(Ray& ray, Intersection& isect) const = 0;
    virtual AABB GetAABBTriangles() const = 0;
    void SetReportBuildProgress(const std::function<void(double, bool)>& f) { report_build_ = f; }

    // Scene::PostConfigure (scene.cpp:47-56) lets emitters fit their shapes to the built
    // scene; the GPU scene derives the environment sphere from its own bounds at Build.
    bool PostConfigure() { return primitives_ != nullptr; }

    // Scene::Intersect (scene.cpp:67-72); environment emitter shapes are handled by the
    // GPU path tracer, so the host query covers the triangles.
    bool Intersect(Ray& ray, Intersection& isect) const { return IntersectTriangles(ray, isect); }

    // Scene::StoreIntersectionFromBarycentricCoords (scene.cpp:101-149)
    void StoreIntersectionFromBarycentricCoords(unsigned int primitiveIndex, unsigned int triangleIndex,
                                                const Ray& ray, const Math::Vec2& b, Intersection& isect) const {
        const Primitive* primitive = primitives_->PrimitiveByIndex(int(primitiveIndex));
        isect.bsdf = primitive->bsdf;
        isect.camera = primitive->camera;
        isect.light = primitive->light;
        const TriangleMesh* mesh = primitive->mesh;
        const Math::Float* P = mesh->Positions();
        const Math::Float* N = mesh->Normals();
        const Math::Float* T = mesh->TexCoords();
        const unsigned int* F = mesh->Faces();
        isect.geom.p = ray.o + ray.d * ray.maxT;
        const unsigned v1 = F[3 * triangleIndex], v2 = F[3 * triangleIndex + 1], v3 = F[3 * triangleIndex + 2];
        auto xyz = [](const Math::Vec4& q) { return Math::Vec3(q.x, q.y, q.z); };
        const Math::Vec3 p1 = xyz(primitive->transform * Math::Vec4(P[3 * v1], P[3 * v1 + 1], P[3 * v1 + 2], 1));
        const Math::Vec3 p2 = xyz(primitive->transform * Math::Vec4(P[3 * v2], P[3 * v2 + 1], P[3 * v2 + 2], 1));
        const Math::Vec3 p3 = xyz(primitive->transform * Math::Vec4(P[3 * v3], P[3 * v3 + 1], P[3 * v3 + 2], 1));
        isect.geom.gn = Math::Normalize(Math::Cross(p2 - p1, p3 - p1));
        auto nrm = [&](unsigned v) {
            return primitive->normalTransform * Math::Vec3(N[3 * v], N[3 * v + 1], N[3 * v + 2]);
        };
        isect.geom.sn = Math::Normalize(nrm(v1) * (1 - b.x - b.y) + nrm(v2) * b.x + nrm(v3) * b.y);
        if (T) {
            isect.geom.uv = Math::Vec2(T[2 * v1] * (1 - b.x - b.y) + T[2 * v2] * b.x + T[2 * v3] * b.y,
                                       T[2 * v1 + 1] * (1 - b.x - b.y) + T[2 * v2 + 1] * b.x + T[2 * v3 + 1] * b.y);
        }
        isect.geom.degenerated = false;
        isect.geom.ComputeTangentSpace();
    }

  protected:
    std::unique_ptr<Primitives> primitives_;
    std::function<void(double, bool)> report_build_;
};

// ---------------------------------------------------------------------------------------
// Rendering

class RenderProcess {
  public:
    virtual ~RenderProcess() {}
};

class SamplingBasedRenderProcess : public RenderProcess {
  public:
    virtual void ProcessSingleSample(const Scene& scene) = 0;
    virtual const Film* GetFilm() const = 0;
};

enum class TerminationMode { Samples, Time };

class Renderer;

class RenderProcessScheduler : public Component {
  public:
    LM_COMPONENT_INTERFACE_DEF("rendersched");
    virtual bool Configure(const ConfigNode& node, const Assets& assets) = 0;
    virtual void SetTerminationMode(TerminationMode mode, double time) = 0;
    virtual bool Render(Renderer& renderer, const Scene& scene) const = 0;
    void SetReportProgress(const std::function<void(double, bool)>& f) { report_ = f; }

  protected:
    std::function<void(double, bool)> report_;
};

class SamplingBasedRenderProcessScheduler : public RenderProcessScheduler {
  public:
    virtual long long NumSamples() const = 0;
};

class Renderer : public Component {
  public:
    LM_COMPONENT_INTERFACE_DEF("renderer");
    virtual std::string Type() const = 0;
    virtual bool Configure(const ConfigNode& node, const Assets& assets, const Scene& scene,
                           const RenderProcessScheduler& sched) = 0;
    virtual bool Preprocess(const Scene& scene, const RenderProcessScheduler& sched) = 0;
    virtual bool Postprocess(const Scene& scene, const RenderProcessScheduler& sched) const = 0;
    virtual RenderProcess* CreateRenderProcess(const Scene& scene, int threadID, int numThreads) = 0;
};

}  // namespace lightmetrica

#endif  // LIGHTMETRICA_AMD_MIRROR_H
