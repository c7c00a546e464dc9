"""Lightmetrica scene files (<nanon version="1.0.dev">) to the GPU path.

Reads the reference's configuration format -- `assets` (raw triangle meshes, BSDFs, films,
cameras, lights), `scene` (a tree of `node`s with `transform`s and asset references, plus
an optional `environment_light`), `renderer` and `render_scheduler` -- and produces the
flattened world-space scene the C ABI takes, with the render settings.

Follows (paths relative to the reference root):
  config.cpp:113-164          root element <nanon>, version "1.0.dev"
  primitives.cpp:96-141       the `root` traversal; the environment light is appended after
                              the area lights met in the traversal
  primitives.cpp:251-431      per-node transform = parent * local, or `global_transform`;
                              camera / light / triangle_mesh + bsdf references; a leaf has no
                              child nodes
  primitives.cpp:433-501      ParseTransform: `matrix` (column-major), else `lookat`, else
                              translate * rotate * scale
  math.transform.inl:25-86    Translate, Rotate (angle in degrees about an axis), Scale
  primitive.h:44              normals use the inverse transpose of the transform
  rawmesh.cpp:65-89           positions, normals (required), texcoords, faces
  diffuse.cpp:67-106, mirror.cpp:60-66, plugin.diffusemirror/diffusemirror.cpp:45-52,
  light.area.cpp, light.env.const.cpp (luminance), perspectivecamera.cpp:97-172 (fovy, film),
  hdrfilm.cpp:80-100 (width, height), pathtrace.cpp:122-135 (rr_depth, max_path_vertices,
  sampler), randomsampler.cpp:42-62 (rng_seed), sched.mt.cpp:70-91 (num_samples).

Only what the GPU path implements is accepted: a BSDF, light, camera or mesh type outside
it raises SceneFormatError naming the element, as the reference fails its Load.
"""
from __future__ import annotations

import time
import xml.etree.ElementTree as ET
from dataclasses import dataclass
from pathlib import Path
from typing import Dict, List, Optional

import numpy as np

from .scenes import (BSDF_DIFFUSE, BSDF_DIFFUSEMIRROR, BSDF_MIRROR, LIGHT_AREA, LIGHT_ENV_CONST, Camera,
                     RenderConfig, Scene, _lights, _materials)

CONFIG_FILE_VERSION = "1.0.dev"

__all__ = ["SceneFormatError", "SceneFile", "load_scene_xml", "parse_transform", "rotate", "translate", "scale",
           "look_at_matrix"]


class SceneFormatError(ValueError):
    pass


@dataclass
class SceneFile:
    scene: Scene
    renderer: str                 # "pt", "pt.direct", "pt.mis" or "raycast"
    scheduler: str                # the render_scheduler type attribute
    primitive_ids: Dict[str, int]  # node id -> primitive index (PrimitiveByID)
    transforms: List[np.ndarray]  # world transform of every primitive (4 x 4, row-major numpy)


# ---- math (math.transform.inl) --------------------------------------------------------

def translate(v) -> np.ndarray:
    m = np.eye(4)
    m[:3, 3] = v
    return m


def rotate(angle_deg: float, axis) -> np.ndarray:
    """Rotation by angle_deg about axis (right-handed): c I + (1 - c) a a^T + s [a]x."""
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    c, s = np.cos(np.radians(angle_deg)), np.sin(np.radians(angle_deg))
    k = np.array([[0.0, -a[2], a[1]], [a[2], 0.0, -a[0]], [-a[1], a[0], 0.0]])
    m = np.eye(4)
    m[:3, :3] = c * np.eye(3) + (1.0 - c) * np.outer(a, a) + s * k
    return m


def scale(v) -> np.ndarray:
    return np.diag([*np.asarray(v, dtype=np.float64).reshape(3), 1.0])


def look_at_matrix(position, center, up) -> np.ndarray:
    """World-to-camera LookAt (math.transform.inl:87-100)."""
    e, c, u = (np.asarray(x, dtype=np.float64) for x in (position, center, up))
    f = c - e
    f /= np.linalg.norm(f)
    s = np.cross(f, u)
    s /= np.linalg.norm(s)
    u2 = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u2, -f
    m[0, 3], m[1, 3], m[2, 3] = -s @ e, -u2 @ e, f @ e
    return m


# ---- element helpers ------------------------------------------------------------------

def _floats(el: Optional[ET.Element], what: str, n: Optional[int] = None) -> np.ndarray:
    if el is None or el.text is None:
        raise SceneFormatError(f"missing '{what}'")
    v = np.array(el.text.split(), dtype=np.float64)
    if n is not None and v.size != n:
        raise SceneFormatError(f"'{what}' needs {n} values, has {v.size}")
    return v


def _child_value(el: ET.Element, name: str, default=None, n: Optional[int] = None):
    c = el.find(name)
    if c is None:
        if default is None:
            raise SceneFormatError(f"<{el.tag} id='{el.get('id', '')}'> misses '{name}'")
        return default
    v = _floats(c, name, n)
    return v if n != 1 else float(v[0])


def parse_transform(el: ET.Element) -> np.ndarray:
    """PrimitivesImpl::ParseTransform (primitives.cpp:433-501)."""
    m = el.find("matrix")
    if m is not None:
        return _floats(m, "matrix", 16).reshape(4, 4).T  # column-major in the file
    la = el.find("lookat")
    if la is not None:
        return look_at_matrix(_floats(la.find("position"), "position", 3), _floats(la.find("center"), "center", 3),
                              _floats(la.find("up"), "up", 3))
    t, r, s = np.eye(4), np.eye(4), np.eye(4)
    if el.find("translate") is not None:
        t = translate(_floats(el.find("translate"), "translate", 3))
    rot = el.find("rotate")
    if rot is not None:
        if rot.find("matrix") is not None:
            r = _floats(rot.find("matrix"), "matrix", 16).reshape(4, 4).T
        elif rot.find("quat") is not None:
            raise SceneFormatError("rotate/quat is not implemented (primitives.cpp:475)")
        else:
            r = rotate(float(_floats(rot.find("angle"), "angle", 1)[0]), _floats(rot.find("axis"), "axis", 3))
    if el.find("scale") is not None:
        s = scale(_floats(el.find("scale"), "scale", 3))
    return t @ r @ s


# ---- loader ---------------------------------------------------------------------------

def _load_assets(assets: ET.Element):
    meshes, bsdfs, films, cameras, lights = {}, {}, {}, {}, {}
    for m in assets.findall("triangle_meshes/triangle_mesh"):
        if m.get("type") != "raw":
            raise SceneFormatError(f"triangle_mesh '{m.get('id')}': type '{m.get('type')}' is not supported "
                                   f"(raw meshes only)")
        P = _floats(m.find("positions"), "positions").reshape(-1, 3)
        N = _floats(m.find("normals"), "normals").reshape(-1, 3)
        F = _floats(m.find("faces"), "faces").astype(np.int64).reshape(-1, 3)
        if N.shape != P.shape:
            raise SceneFormatError(f"triangle_mesh '{m.get('id')}': one normal per position is required")
        if F.size and (F.min() < 0 or F.max() >= len(P)):
            raise SceneFormatError(f"triangle_mesh '{m.get('id')}': face index out of range")
        meshes[m.get("id")] = (P, N, F)
    for b in assets.findall("bsdfs/bsdf"):
        t = b.get("type")
        if t == "diffuse":
            dr = b.find("diffuse_reflectance")
            if dr is None:
                raise SceneFormatError(f"bsdf '{b.get('id')}': missing 'diffuse_reflectance'")
            if dr.find("texture") is not None:
                raise SceneFormatError(f"bsdf '{b.get('id')}': textures are not supported on the GPU path")
            bsdfs[b.get("id")] = (BSDF_DIFFUSE, _floats(dr.find("color"), "color", 3))
        elif t == "mirror":
            bsdfs[b.get("id")] = (BSDF_MIRROR, _child_value(b, "specular_reflectance", np.ones(3), 3))
        elif t == "plugin.diffusemirror":
            bsdfs[b.get("id")] = (BSDF_DIFFUSEMIRROR, _child_value(b, "diffuse_reflectance", np.ones(3), 3))
        else:
            raise SceneFormatError(f"bsdf '{b.get('id')}': type '{t}' is not supported on the GPU path")
    for f in assets.findall("films/film"):
        films[f.get("id")] = (int(_child_value(f, "width", n=1)), int(_child_value(f, "height", n=1)))
    for c in assets.findall("cameras/camera"):
        if c.get("type") != "perspective":
            raise SceneFormatError(f"camera '{c.get('id')}': type '{c.get('type')}' is not supported")
        fr = c.find("film")
        if fr is None or fr.get("ref") not in films:
            raise SceneFormatError(f"camera '{c.get('id')}': missing or unknown film reference")
        cameras[c.get("id")] = (_child_value(c, "fovy", n=1), films[fr.get("ref")])
    for l in assets.findall("lights/light"):
        t = l.get("type")
        if t not in ("area", "env.const"):
            raise SceneFormatError(f"light '{l.get('id')}': type '{t}' is not supported on the GPU path")
        lights[l.get("id")] = (LIGHT_AREA if t == "area" else LIGHT_ENV_CONST, _child_value(l, "luminance", n=3))
    return meshes, bsdfs, films, cameras, lights


def _ref(node: ET.Element, tag: str, table: dict, what: str):
    r = node.find(tag)
    if r is None:
        return None
    key = r.get("ref")
    if key not in table:
        raise SceneFormatError(f"unknown {what} reference '{key}'")
    return key


def load_scene_xml(source, num_samples: Optional[int] = None) -> SceneFile:
    """Parse a scene file (path or XML text) into the GPU scene and its render settings."""
    text = Path(source).read_text() if not str(source).lstrip().startswith("<") else str(source)
    root = ET.fromstring(text)
    if root.tag != "nanon":
        raise SceneFormatError("missing <nanon> element")
    if root.get("version") != CONFIG_FILE_VERSION:
        raise SceneFormatError(f"different version : {root.get('version')} ( expected : {CONFIG_FILE_VERSION} )")
    assets_el, scene_el = root.find("assets"), root.find("scene")
    if assets_el is None or scene_el is None:
        raise SceneFormatError("missing 'assets' or 'scene' element")
    meshes, bsdfs, films, cameras, lights = _load_assets(assets_el)
    top = scene_el.find("root")
    if top is None:
        raise SceneFormatError("missing 'root' node")

    P_all, N_all, F_all, M_all, L_all = [], [], [], [], []
    mat_index: Dict[str, int] = {}
    mat_entries = []
    light_order: List[str] = []  # area lights in traversal order (primitives.cpp:300)
    camera_desc = None
    ids: Dict[str, int] = {}
    transforms: List[np.ndarray] = []
    nv = 0

    def traverse(node: ET.Element, parent: np.ndarray):
        nonlocal nv, camera_desc
        local = parse_transform(node.find("transform")) if node.find("transform") is not None else np.eye(4)
        g = node.find("global_transform")
        world = parse_transform(g) if g is not None else parent @ local
        cam_key = _ref(node, "camera", cameras, "camera")
        light_key = _ref(node, "light", lights, "light")
        mesh_key = _ref(node, "triangle_mesh", meshes, "triangle_mesh")
        if cam_key and light_key:
            raise SceneFormatError("'camera' and 'light' elements cannot be used in the same time")
        is_prim = bool(cam_key or light_key or mesh_key)
        if cam_key and camera_desc is None:  # the first camera is the main one
            camera_desc = (cameras[cam_key], world)
        if light_key and light_key not in light_order:
            light_order.append(light_key)
        if mesh_key:
            bsdf_key = _ref(node, "bsdf", bsdfs, "bsdf")
            if bsdf_key is None:
                raise SceneFormatError("missing 'bsdf' element")
            if bsdf_key not in mat_index:
                mat_index[bsdf_key] = len(mat_entries)
                mat_entries.append(bsdfs[bsdf_key])
            P, N, F = meshes[mesh_key]
            Pw = (np.c_[P, np.ones(len(P))] @ world.T)[:, :3]
            Nw = N @ np.linalg.inv(world[:3, :3])  # (M^-1)^T n, row vectors
            P_all.append(Pw)
            N_all.append(Nw)
            F_all.append(F + nv)
            M_all.append(np.full(len(F), mat_index[bsdf_key]))
            L_all.append(np.full(len(F), -1 if light_key is None else light_order.index(light_key)))
            nv += len(P)
        if is_prim:
            transforms.append(world)
            if node.get("id"):
                if node.get("id") in ids:
                    raise SceneFormatError(f"ID '{node.get('id')}' for the node is already used")
                ids[node.get("id")] = len(transforms) - 1
        children = node.findall("node")
        if children and is_prim:
            raise SceneFormatError("Leaf node cannot have children")
        for ch in children:
            traverse(ch, world)

    traverse(top, np.eye(4))
    env = scene_el.find("environment_light")
    light_keys = list(light_order)
    if env is not None:
        key = env.get("ref")
        if key not in lights or lights[key][0] != LIGHT_ENV_CONST:
            raise SceneFormatError("The light referenced by 'environment_light' is not an environment light")
        light_keys.append(key)
    if camera_desc is None:
        raise SceneFormatError("missing 'camera' in the scene")
    if not F_all:
        raise SceneFormatError("the scene has no triangle meshes")
    (fovy, (W, H)), view = camera_desc

    renderer_el = root.find("renderer")
    rtype = renderer_el.get("type") if renderer_el is not None else "pt"
    rtype = rtype[:-4] if rtype.endswith(".hip") else rtype
    if rtype not in ("pt", "pt.direct", "pt.mis", "raycast"):
        raise SceneFormatError(f"renderer '{rtype}' is not supported on the GPU path")
    rr_depth = int(_child_value(renderer_el, "rr_depth", 1.0, 1)) if renderer_el is not None else 1
    max_pv = int(_child_value(renderer_el, "max_path_vertices", -1.0, 1)) if renderer_el is not None else -1
    seed = -1
    sampler = renderer_el.find("sampler") if renderer_el is not None else None
    if sampler is not None and sampler.find("rng_seed") is not None:
        seed = int(_floats(sampler.find("rng_seed"), "rng_seed", 1)[0])
    if seed < 0:
        seed = int(time.time())  # randomsampler.cpp:54-57
    sched_el = root.find("render_scheduler")
    stype = sched_el.get("type") if sched_el is not None else "mt"
    ns = int(_child_value(sched_el, "num_samples", 1.0, 1)) if sched_el is not None else 1
    if num_samples is not None:
        ns = int(num_samples)

    mats = _materials([(t, tuple(r)) for t, r in mat_entries])
    lts = _lights([(lights[k][0], tuple(lights[k][1])) for k in light_keys])
    cam = Camera(view=view.T.reshape(-1).astype(np.float32), fovy=float(fovy), width=W, height=H)
    rc = RenderConfig(num_samples=ns, max_path_vertices=max_pv, rr_depth=rr_depth, seed=seed)
    sc = Scene("xml", np.concatenate(P_all).astype(np.float32), np.concatenate(N_all).astype(np.float32),
               np.concatenate(F_all).astype(np.uint32), np.concatenate(M_all).astype(np.uint32),
               np.concatenate(L_all).astype(np.int32), mats, lts, cam, rc)
    return SceneFile(sc, rtype, stype, ids, transforms)
