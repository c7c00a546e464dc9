"""Synthetic scenes and ray batches of the benchmark configurations.

Every random quantity comes from Philox4x32-10 (Salmon et al., SC'11), so the CPU
restatement, the GPU kernels and these generators see identical inputs.

Configurations (world-space triangle soups, per-vertex normals, per-triangle
material and light ids, the way Primitives::Traverse flattens a scene,
src/liblightmetrica/primitives.cpp:251-418):

* ``room_scene``   -- closed box room of 5 quads, two rotated boxes, one 0.25 x 0.25
  ceiling light (17, 12, 4); pinhole camera, 60 degree fov, 256 x 256, 16 spp,
  max depth 8 (max_path_vertices 9), Russian roulette from depth 3.
* ``traversal_scene`` / ``traversal_rays`` -- 1M triangles with centroids uniform in
  the unit cube and vertices offset by N(0, 0.01); 16M rays with origins uniform in
  [-0.5, 1.5]^3 and directions uniform on the sphere.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

_MASK32 = np.uint64(0xFFFFFFFF)

MATERIAL_DTYPE = np.dtype([("type", "<u4"), ("R", "<f4", (3,))])
LIGHT_DTYPE = np.dtype([("type", "<u4"), ("Le", "<f4", (3,))])

BSDF_DIFFUSE, BSDF_MIRROR, BSDF_DIFFUSEMIRROR = 0, 1, 2
LIGHT_AREA, LIGHT_ENV_CONST = 0, 1

# Stream tags of the generators (counter word 3); the path-sample streams use 'LMPT'.
TAG_GEOMETRY = 0x47454F4D  # 'GEOM'
TAG_RAYS = 0x52415953      # 'RAYS'


def philox4x32_10(c0, c1, c2, c3, k0: int, k1: int):
    """Vectorised Philox4x32-10.  Counters are uint32 arrays (broadcast), key two uint32."""
    c0, c1, c2, c3 = (np.asarray(x, dtype=np.uint64) & _MASK32 for x in (c0, c1, c2, c3))
    c0, c1, c2, c3 = np.broadcast_arrays(c0, c1, c2, c3)
    k0 = np.uint64(k0 & 0xFFFFFFFF)
    k1 = np.uint64(k1 & 0xFFFFFFFF)
    m0, m1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
    w0, w1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
    for _ in range(10):
        p0 = m0 * c0
        p1 = m1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK32
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
        k0 = (k0 + w0) & _MASK32
        k1 = (k1 + w1) & _MASK32
    return tuple(x.astype(np.uint32) for x in (c0, c1, c2, c3))


def uniform_blocks(seed: int, n: int, blocks: int, tag: int) -> np.ndarray:
    """(n, 4*blocks) float64 uniforms in [0, 1) with 24-bit resolution; row i uses counters (b, i, 0, tag)."""
    idx = np.arange(n, dtype=np.uint64)
    out = np.empty((n, 4 * blocks), dtype=np.float64)
    for b in range(blocks):
        words = philox4x32_10(np.uint64(b), idx, np.uint64(0), np.uint64(tag), seed & 0xFFFFFFFF, seed >> 32)
        for k in range(4):
            out[:, 4 * b + k] = (words[k] >> np.uint32(8)).astype(np.float64) * (1.0 / 16777216.0)
    return out


def sample_stream_u32(seed: int, sample: int, count: int) -> np.ndarray:
    """First `count` draws of one path sample's stream (seed, sample) -- the kernels' SampleStream."""
    blocks = (count + 3) // 4
    words = philox4x32_10(np.arange(blocks, dtype=np.uint64), np.uint64(sample & 0xFFFFFFFF),
                          np.uint64(sample >> 32), np.uint64(0x4C4D5054), seed & 0xFFFFFFFF, seed >> 32)
    return np.stack(words, axis=1).reshape(-1)[:count]


def look_at(eye, center, up) -> np.ndarray:
    """Math::LookAt (include/lightmetrica/math.transform.inl:87-100) as a column-major 4x4 (float64)."""
    eye, center, up = (np.asarray(v, dtype=np.float64) for v in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    u = up / np.linalg.norm(up)
    s = np.cross(f, u)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.zeros((4, 4))  # m[col, row]
    m[0] = [s[0], u[0], -f[0], 0]
    m[1] = [s[1], u[1], -f[1], 0]
    m[2] = [s[2], u[2], -f[2], 0]
    m[3] = [-s @ eye, -u @ eye, f @ eye, 1]
    return m.reshape(-1)


@dataclass
class Camera:
    view: np.ndarray  # 16 float32, column-major world->camera
    fovy: float
    width: int
    height: int


@dataclass
class RenderConfig:
    num_samples: int
    max_path_vertices: int
    rr_depth: int
    seed: int = 1


@dataclass
class Scene:
    name: str
    positions: np.ndarray
    normals: np.ndarray
    faces: np.ndarray
    material_ids: Optional[np.ndarray] = None
    light_ids: Optional[np.ndarray] = None
    materials: Optional[np.ndarray] = None
    lights: Optional[np.ndarray] = None
    camera: Optional[Camera] = None
    render: Optional[RenderConfig] = None
    meta: dict = field(default_factory=dict)

    @property
    def num_triangles(self) -> int:
        return int(self.faces.shape[0])


class _Builder:
    def __init__(self):
        self.P, self.N, self.F, self.M, self.L = [], [], [], [], []
        self.nv = 0

    def quad(self, corner, a, b, mat: int, light: int = -1):
        corner, a, b = (np.asarray(v, dtype=np.float64) for v in (corner, a, b))
        n = np.cross(a, b)
        n /= np.linalg.norm(n)
        self.P += [corner, corner + a, corner + a + b, corner + b]
        self.N += [n] * 4
        v = self.nv
        self.F += [(v, v + 1, v + 2), (v, v + 2, v + 3)]
        self.M += [mat, mat]
        self.L += [light, light]
        self.nv += 4

    def box(self, center, size, angle_deg: float, mat: int):
        h = np.asarray(size, dtype=np.float64) / 2.0
        c, s = np.cos(np.radians(angle_deg)), np.sin(np.radians(angle_deg))
        R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
        faces = [
            ((h[0], -h[1], -h[2]), (0, 2 * h[1], 0), (0, 0, 2 * h[2])),
            ((-h[0], -h[1], -h[2]), (0, 0, 2 * h[2]), (0, 2 * h[1], 0)),
            ((-h[0], h[1], -h[2]), (0, 0, 2 * h[2]), (2 * h[0], 0, 0)),
            ((-h[0], -h[1], -h[2]), (2 * h[0], 0, 0), (0, 0, 2 * h[2])),
            ((-h[0], -h[1], h[2]), (2 * h[0], 0, 0), (0, 2 * h[1], 0)),
            ((-h[0], -h[1], -h[2]), (0, 2 * h[1], 0), (2 * h[0], 0, 0)),
        ]
        ctr = np.asarray(center, dtype=np.float64)
        for corner, a, b in faces:
            self.quad(R @ np.asarray(corner) + ctr, R @ np.asarray(a), R @ np.asarray(b), mat)

    def arrays(self):
        return (np.asarray(self.P, dtype=np.float32), np.asarray(self.N, dtype=np.float32),
                np.asarray(self.F, dtype=np.uint32), np.asarray(self.M, dtype=np.uint32),
                np.asarray(self.L, dtype=np.int32))


def _materials(entries) -> np.ndarray:
    m = np.zeros(len(entries), dtype=MATERIAL_DTYPE)
    for i, (t, r) in enumerate(entries):
        m[i]["type"] = t
        m[i]["R"] = r
    return m


def _lights(entries) -> np.ndarray:
    m = np.zeros(len(entries), dtype=LIGHT_DTYPE)
    for i, (t, le) in enumerate(entries):
        m[i]["type"] = t
        m[i]["Le"] = le
    return m


def room_scene(width: int = 256, height: int = 256, spp: int = 16, seed: int = 1) -> Scene:
    """The closed box room with two rotated boxes and one ceiling area light."""
    b = _Builder()
    WHITE, RED, GREEN, BLACK = 0, 1, 2, 3
    b.quad((0, 0, 0), (0, 0, 1), (1, 0, 0), WHITE)   # floor y = 0
    b.quad((0, 1, 0), (1, 0, 0), (0, 0, 1), WHITE)   # ceiling y = 1
    b.quad((0, 0, 0), (1, 0, 0), (0, 1, 0), WHITE)   # back z = 0
    b.quad((0, 0, 0), (0, 1, 0), (0, 0, 1), RED)     # left x = 0
    b.quad((1, 0, 0), (0, 0, 1), (0, 1, 0), GREEN)   # right x = 1
    b.box((0.35, 0.15, 0.35), (0.3, 0.3, 0.3), 18.0, WHITE)
    b.box((0.68, 0.30, 0.62), (0.3, 0.6, 0.3), -20.0, WHITE)
    b.quad((0.375, 0.999, 0.375), (0.25, 0, 0), (0, 0, 0.25), BLACK, light=0)
    P, N, F, M, L = b.arrays()
    mats = _materials([(BSDF_DIFFUSE, (0.75, 0.75, 0.75)), (BSDF_DIFFUSE, (0.63, 0.06, 0.04)),
                       (BSDF_DIFFUSE, (0.14, 0.45, 0.09)), (BSDF_DIFFUSE, (0.0, 0.0, 0.0))])
    lights = _lights([(LIGHT_AREA, (17.0, 12.0, 4.0))])
    view = look_at((0.5, 0.5, 1.95), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0)).astype(np.float32)
    cam = Camera(view=view, fovy=60.0, width=width, height=height)
    rc = RenderConfig(num_samples=width * height * spp, max_path_vertices=9, rr_depth=3, seed=seed)
    return Scene("room", P, N, F, M, L, mats, lights, cam, rc)


def _axis_boxes(centers: np.ndarray, half: np.ndarray, subdiv: int, top_only_grid: bool = False,
                skip_bottom: bool = False):
    """Axis-aligned boxes, outward normals, each face split into subdiv x subdiv quads
    (with top_only_grid only the +y face is split; the sides are single quads).
    Returns P (n*V, 3), N (n*V, 3), F (n*T, 3) and the number of triangles per box."""
    n = centers.shape[0]
    quads = []  # per quad: corner (n, 3), du (n, 3), dv (n, 3), normal (3,)
    for k in range(3):
        i, j = (k + 1) % 3, (k + 2) % 3
        for sgn in (1.0, -1.0):
            if skip_bottom and k == 1 and sgn < 0:
                continue
            a, b = (i, j) if sgn > 0 else (j, i)  # e_a x e_b = sgn e_k
            s = subdiv if (not top_only_grid or (k == 1 and sgn > 0)) else 1
            for ia in range(s):
                for ib in range(s):
                    corner = centers.copy()
                    corner[:, k] += sgn * half[:, k]
                    corner[:, a] += -half[:, a] + 2.0 * half[:, a] * ia / s
                    corner[:, b] += -half[:, b] + 2.0 * half[:, b] * ib / s
                    du = np.zeros_like(centers)
                    dv = np.zeros_like(centers)
                    du[:, a] = 2.0 * half[:, a] / s
                    dv[:, b] = 2.0 * half[:, b] / s
                    nrm = np.zeros(3)
                    nrm[k] = sgn
                    quads.append((corner, du, dv, nrm))
    q = len(quads)
    P = np.empty((n, q, 4, 3), dtype=np.float64)
    Nn = np.empty((n, q, 4, 3), dtype=np.float64)
    for t, (corner, du, dv, nrm) in enumerate(quads):
        P[:, t, 0] = corner
        P[:, t, 1] = corner + du
        P[:, t, 2] = corner + du + dv
        P[:, t, 3] = corner + dv
        Nn[:, t, :] = nrm
    base = (np.arange(n * q, dtype=np.int64) * 4)[:, None]
    F = np.concatenate([base + np.array([0, 1, 2]), base + np.array([0, 2, 3])], axis=1).reshape(-1, 3)
    return P.reshape(-1, 3), Nn.reshape(-1, 3), F, 2 * q


def gradient_scene(width: int = 512, height: int = 512, spp: int = 32, num_boxes: int = 2000, seed: int = 0xC4,
                   box_albedo=None, max_path_vertices: int = 9, rr_depth: int = 9) -> Scene:
    """The room of room_scene plus `num_boxes` axis-aligned boxes (edges U[0.01, 0.05],
    centres in [0.05, 0.95] x [0, 0.9] x [0.05, 0.95], each face a 2 x 2 grid: 48 triangles)
    with one diffuse material per box, and four 0.2 x 0.2 ceiling lights.  Materials:
    0 white, 1 red, 2 green, 3 black (lights), 4 + i box i.  rr_depth = max_path_vertices
    disables Russian roulette, so with common sample streams the image is a polynomial in
    the reflectances and emissions."""
    b = _Builder()
    WHITE, RED, GREEN, BLACK = 0, 1, 2, 3
    b.quad((0, 0, 0), (0, 0, 1), (1, 0, 0), WHITE)
    b.quad((0, 1, 0), (1, 0, 0), (0, 0, 1), WHITE)
    b.quad((0, 0, 0), (1, 0, 0), (0, 1, 0), WHITE)
    b.quad((0, 0, 0), (0, 1, 0), (0, 0, 1), RED)
    b.quad((1, 0, 0), (0, 0, 1), (0, 1, 0), GREEN)
    b.box((0.35, 0.15, 0.35), (0.3, 0.3, 0.3), 18.0, WHITE)
    b.box((0.68, 0.30, 0.62), (0.3, 0.6, 0.3), -20.0, WHITE)
    for li, (cx, cz) in enumerate(((0.3, 0.3), (0.7, 0.3), (0.3, 0.7), (0.7, 0.7))):
        b.quad((cx - 0.1, 0.999, cz - 0.1), (0.2, 0, 0), (0, 0, 0.2), BLACK, light=li)
    P0, N0, F0, M0, L0 = b.arrays()
    u = uniform_blocks(seed, num_boxes, 2, TAG_GEOMETRY)
    half = (0.01 + 0.04 * u[:, 0:3]) / 2.0
    ctr = np.stack([0.05 + 0.9 * u[:, 3], 0.9 * u[:, 4], 0.05 + 0.9 * u[:, 5]], axis=1)
    P1, N1, F1, tpb = _axis_boxes(ctr, half, 2)
    P = np.concatenate([P0, P1.astype(np.float32)])
    Nn = np.concatenate([N0, N1.astype(np.float32)])
    F = np.concatenate([F0, (F1 + P0.shape[0]).astype(np.uint32)])
    M = np.concatenate([M0, np.repeat(4 + np.arange(num_boxes, dtype=np.uint32), tpb)])
    L = np.concatenate([L0, np.full(num_boxes * tpb, -1, dtype=np.int32)])
    if box_albedo is None:
        box_albedo = np.full((num_boxes, 3), 0.5)
    entries = [(BSDF_DIFFUSE, (0.75, 0.75, 0.75)), (BSDF_DIFFUSE, (0.63, 0.06, 0.04)),
               (BSDF_DIFFUSE, (0.14, 0.45, 0.09)), (BSDF_DIFFUSE, (0.0, 0.0, 0.0))]
    mats = _materials(entries + [(BSDF_DIFFUSE, tuple(r)) for r in np.asarray(box_albedo, dtype=np.float64)])
    lights = _lights([(LIGHT_AREA, (8.5, 6.0, 2.0))] * 4)
    view = look_at((0.5, 0.5, 1.95), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0)).astype(np.float32)
    cam = Camera(view=view, fovy=60.0, width=width, height=height)
    rc = RenderConfig(num_samples=width * height * spp, max_path_vertices=max_path_vertices, rr_depth=rr_depth,
                      seed=1)
    return Scene("gradient", P, Nn, F, M, L, mats, lights, cam, rc,
                 meta={"box_materials": 4 + np.arange(num_boxes), "seed": seed})


def hidden_box_albedo(num_boxes: int = 2000, seed: int = 0xC4) -> np.ndarray:
    """The hidden per-box reflectances of the gradient workload's target image: U[0.1, 0.9]^3."""
    return 0.1 + 0.8 * uniform_blocks(seed ^ 0x7A46, num_boxes, 1, TAG_GEOMETRY)[:, 0:3]


def _gradient_noise(x: np.ndarray, z: np.ndarray, freq: int, seed: int) -> np.ndarray:
    """2-D gradient noise (Perlin, quintic fade) on a freq x freq lattice over [0, 1]^2,
    unit gradients at angles drawn from Philox4x32-10."""
    ang = 2.0 * np.pi * uniform_blocks(seed, (freq + 1) * (freq + 1), 1, TAG_GEOMETRY)[:, 0]
    gx, gz = np.cos(ang).reshape(freq + 1, freq + 1), np.sin(ang).reshape(freq + 1, freq + 1)
    fx, fz = x * freq, z * freq
    ix = np.minimum(np.floor(fx).astype(np.int64), freq - 1)
    iz = np.minimum(np.floor(fz).astype(np.int64), freq - 1)
    tx, tz = fx - ix, fz - iz

    def corner(dx, dz):
        return gx[ix + dx, iz + dz] * (tx - dx) + gz[ix + dx, iz + dz] * (tz - dz)

    fade = lambda t: t * t * t * (t * (t * 6.0 - 15.0) + 10.0)
    u, v = fade(tx), fade(tz)
    a = corner(0, 0) + u * (corner(1, 0) - corner(0, 0))
    b = corner(0, 1) + u * (corner(1, 1) - corner(0, 1))
    return a + v * (b - a)


def terrain_scene(width: int = 1920, height: int = 1080, spp: int = 64, grid: int = 1024, seed: int = 0xC3,
                  mirror_fraction: float = 0.4) -> Scene:
    """Heightfield terrain: (grid + 1)^2 vertices over [0, 1]^2 (x, z), height 0.2 * fBm
    (6 octaves of gradient noise from 4 cells, lacunarity 2, gain 0.5), 2 grid^2 triangles
    with interpolated normals.  Per triangle: a mirror (R = 1) with probability 0.4, else a
    diffuse material of its own with reflectance U[0.1, 0.9]^3.  A 1 x 1 quad light at
    y = 1.5 facing down (Le = 10) and a constant environment (Le = 0.5), in the reference's
    light order (area lights, then the environment light)."""
    n = grid + 1
    c = np.arange(n, dtype=np.float64) / grid
    X, Z = np.meshgrid(c, c, indexing="ij")
    Y = np.zeros_like(X)
    amp, freq = 1.0, 4
    for o in range(6):
        Y += amp * _gradient_noise(X, Z, freq, seed + 101 * o)
        amp *= 0.5
        freq *= 2
    Y *= 0.2
    # Normals from central differences of the height field.
    dx = np.gradient(Y, c, axis=0)
    dz = np.gradient(Y, c, axis=1)
    Nv = np.stack([-dx, np.ones_like(Y), -dz], axis=-1)
    Nv /= np.linalg.norm(Nv, axis=-1, keepdims=True)
    P = np.stack([X, Y, Z], axis=-1).reshape(-1, 3)
    idx = np.arange(n * n, dtype=np.int64).reshape(n, n)
    v00, v10 = idx[:-1, :-1].ravel(), idx[1:, :-1].ravel()
    v01, v11 = idx[:-1, 1:].ravel(), idx[1:, 1:].ravel()
    F = np.empty((2 * grid * grid, 3), dtype=np.int64)
    F[0::2] = np.stack([v00, v01, v11], axis=1)
    F[1::2] = np.stack([v00, v11, v10], axis=1)
    nt = F.shape[0]
    u = uniform_blocks(seed ^ 0x5EED, nt, 1, TAG_GEOMETRY)
    mirror = u[:, 0] < mirror_fraction
    # Material 0: black (light), 1: the shared mirror, 2 + k: the k-th diffuse triangle.
    M = np.empty(nt, dtype=np.uint32)
    M[mirror] = 1
    nd = int((~mirror).sum())
    M[~mirror] = 2 + np.arange(nd, dtype=np.uint32)
    mats = np.zeros(2 + nd, dtype=MATERIAL_DTYPE)
    mats[0]["type"], mats[1]["type"] = BSDF_DIFFUSE, BSDF_MIRROR
    mats[1]["R"] = (1.0, 1.0, 1.0)
    mats["type"][2:] = BSDF_DIFFUSE
    mats["R"][2:] = 0.1 + 0.8 * u[~mirror, 1:4]
    # The light quad: corner (0, 1.5, 0), spanning x and z, normal -y.
    lp = np.array([[0, 1.5, 0], [1, 1.5, 0], [1, 1.5, 1], [0, 1.5, 1]], dtype=np.float64)
    base = P.shape[0]
    P = np.concatenate([P, lp]).astype(np.float32)
    Nn = np.concatenate([Nv.reshape(-1, 3), np.tile([0.0, -1.0, 0.0], (4, 1))]).astype(np.float32)
    F = np.concatenate([F, [[base, base + 1, base + 2], [base, base + 2, base + 3]]]).astype(np.uint32)
    M = np.concatenate([M, [0, 0]]).astype(np.uint32)
    L = np.concatenate([np.full(nt, -1), [0, 0]]).astype(np.int32)
    lights = _lights([(LIGHT_AREA, (10.0, 10.0, 10.0)), (LIGHT_ENV_CONST, (0.5, 0.5, 0.5))])
    view = look_at((0.5, 0.6, 1.6), (0.5, 0.1, 0.5), (0.0, 1.0, 0.0)).astype(np.float32)
    cam = Camera(view=view, fovy=45.0, width=width, height=height)
    rc = RenderConfig(num_samples=width * height * spp, max_path_vertices=17, rr_depth=3, seed=1)
    return Scene("terrain", P, Nn, F, M, L, mats, lights, cam, rc, meta={"seed": seed, "grid": grid})


def city_scene(width: int = 3840, height: int = 2160, spp: int = 256, cells: int = 1000, seed: int = 0xC5,
               groups: int = 10_000, num_lights: int = 16) -> Scene:
    """Procedural city: one axis-aligned box per unit cell of a cells x cells grid, footprint
    0.8 x 0.8, height exp(N(0, 0.8^2)); 16 triangles per box (4 sides, a 2 x 2 top, no
    bottom).  Box i takes albedo group (2654435761 i mod 2^32) mod `groups`, each group a
    diffuse reflectance U[0.1, 0.9]^3.  A ground quad (reflectance 0.5) and a 4 x 4 grid of
    40 x 40 quad lights at y = 60 facing down.  Materials: 0 .. groups-1 the groups, then
    the ground, then black (lights)."""
    n = cells * cells
    u = uniform_blocks(seed, n, 1, TAG_GEOMETRY)
    r = np.sqrt(-2.0 * np.log(1.0 - u[:, 0]))
    hgt = np.exp(0.8 * r * np.cos(2.0 * np.pi * u[:, 1]))
    ii, jj = np.meshgrid(np.arange(cells, dtype=np.float64), np.arange(cells, dtype=np.float64), indexing="ij")
    ctr = np.stack([ii.ravel() + 0.5, hgt / 2.0, jj.ravel() + 0.5], axis=1)
    half = np.stack([np.full(n, 0.4), hgt / 2.0, np.full(n, 0.4)], axis=1)
    P, Nn, F, tpb = _axis_boxes(ctr, half, 2, top_only_grid=True, skip_bottom=True)
    grp = ((np.arange(n, dtype=np.uint64) * np.uint64(2654435761)) & _MASK32) % np.uint64(groups)
    M = np.repeat(grp.astype(np.uint32), tpb)
    ground, black = groups, groups + 1
    ext = float(cells)
    gq = np.array([[0, 0, 0], [0, 0, ext], [ext, 0, ext], [ext, 0, 0]], dtype=np.float64)  # normal +y
    lq = []
    k = int(round(np.sqrt(num_lights)))
    for a in range(k):
        for b in range(k):
            cx, cz = (a + 0.5) * ext / k, (b + 0.5) * ext / k
            lq.append([[cx - 20, 60, cz - 20], [cx + 20, 60, cz - 20], [cx + 20, 60, cz + 20], [cx - 20, 60, cz + 20]])
    extra = np.concatenate([gq[None], np.asarray(lq, dtype=np.float64)]).reshape(-1, 3)
    base = P.shape[0]
    quads = (base + 4 * np.arange(1 + len(lq)))[:, None]
    Fx = np.concatenate([quads + np.array([0, 1, 2]), quads + np.array([0, 2, 3])], axis=1).reshape(-1, 3)
    Nx = np.concatenate([np.tile([0.0, 1.0, 0.0], (4, 1)), np.tile([0.0, -1.0, 0.0], (4 * len(lq), 1))])
    P = np.concatenate([P, extra]).astype(np.float32)
    Nn = np.concatenate([Nn, Nx]).astype(np.float32)
    F = np.concatenate([F, Fx]).astype(np.uint32)
    M = np.concatenate([M, [ground, ground], np.full(2 * len(lq), black)]).astype(np.uint32)
    L = np.concatenate([np.full(n * tpb + 2, -1), np.repeat(np.arange(len(lq)), 2)]).astype(np.int32)
    alb = 0.1 + 0.8 * uniform_blocks(seed ^ 0xA1BE, groups, 1, TAG_GEOMETRY)[:, 0:3]
    mats = np.zeros(groups + 2, dtype=MATERIAL_DTYPE)
    mats["type"] = BSDF_DIFFUSE
    mats["R"][:groups] = alb
    mats["R"][ground] = (0.5, 0.5, 0.5)
    lights = _lights([(LIGHT_AREA, (5.0, 5.0, 5.0))] * len(lq))
    view = look_at((ext / 2, 120.0, ext * 1.08), (ext / 2, 0.0, ext * 0.45), (0.0, 1.0, 0.0)).astype(np.float32)
    cam = Camera(view=view, fovy=40.0, width=width, height=height)
    rc = RenderConfig(num_samples=width * height * spp, max_path_vertices=13, rr_depth=3, seed=1)
    return Scene("city", P, Nn, F, M, L, mats, lights, cam, rc, meta={"seed": seed, "cells": cells})


def traversal_scene(num_triangles: int = 1_000_000, seed: int = 0xC2A, sigma: float = 0.01) -> Scene:
    """Triangles with centroids uniform in [0,1]^3 and vertices at centroid + N(0, sigma^2 I)."""
    u = uniform_blocks(seed, num_triangles, 4, TAG_GEOMETRY)  # 16 uniforms per triangle
    c = u[:, 0:3]
    # Box-Muller on uniforms (0, 1]: 5 pairs -> 10 normals, 9 used.
    u1 = 1.0 - u[:, 3:8]
    u2 = u[:, 8:13]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)], axis=1)[:, :9]
    v = c[:, None, :] + sigma * z.reshape(-1, 3, 3)
    P = v.reshape(-1, 3).astype(np.float32)
    F = np.arange(3 * num_triangles, dtype=np.uint32).reshape(-1, 3)
    e1 = P[F[:, 1]].astype(np.float64) - P[F[:, 0]]
    e2 = P[F[:, 2]].astype(np.float64) - P[F[:, 0]]
    n = np.cross(e1, e2)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    N = np.repeat(n, 3, axis=0).astype(np.float32)
    return Scene("traversal", P, N, F, meta={"seed": seed, "sigma": sigma})


@dataclass
class Rays:
    o: np.ndarray     # (n, 3) float32
    d: np.ndarray     # (n, 3) float32
    tmin: np.ndarray  # (n,) float32
    tmax: np.ndarray  # (n,) float32

    def __len__(self) -> int:
        return int(self.o.shape[0])

    def slice(self, a: int, b: int) -> "Rays":
        return Rays(self.o[a:b], self.d[a:b], self.tmin[a:b], self.tmax[a:b])


FLT_MAX = np.float32(3.4028234663852886e38)


def traversal_rays(num_rays: int = 16_000_000, seed: int = 0xC2B, chunk: int = 1 << 22) -> Rays:
    """Origins uniform in [-0.5, 1.5]^3, directions UniformSampleSphere (math.stats.inl:105-112)."""
    o = np.empty((num_rays, 3), dtype=np.float32)
    d = np.empty((num_rays, 3), dtype=np.float32)
    for a in range(0, num_rays, chunk):
        b = min(num_rays, a + chunk)
        idx = np.arange(a, b, dtype=np.uint64)
        words = [philox4x32_10(np.uint64(0), idx, np.uint64(0), np.uint64(TAG_RAYS), seed & 0xFFFFFFFF, seed >> 32),
                 philox4x32_10(np.uint64(1), idx, np.uint64(0), np.uint64(TAG_RAYS), seed & 0xFFFFFFFF, seed >> 32)]
        u = [(w >> np.uint32(8)).astype(np.float64) * (1.0 / 16777216.0) for w in (*words[0], *words[1])]
        o[a:b, 0] = -0.5 + 2.0 * u[0]
        o[a:b, 1] = -0.5 + 2.0 * u[1]
        o[a:b, 2] = -0.5 + 2.0 * u[2]
        z = 1.0 - 2.0 * u[4]
        r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
        phi = 2.0 * np.pi * u[5]
        d[a:b, 0] = r * np.cos(phi)
        d[a:b, 1] = r * np.sin(phi)
        d[a:b, 2] = z
    return Rays(o, d, np.zeros(num_rays, dtype=np.float32), np.full(num_rays, FLT_MAX, dtype=np.float32))


def stub_simple_mesh() -> Scene:
    """StubTriangleMesh_Simple (include/lightmetrica.test/stub.trianglemesh.h:58-122): two unit squares at z=0, -1."""
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                  [0, 0, -1], [1, 0, -1], [1, 1, -1], [0, 1, -1]], dtype=np.float32)
    N = np.tile(np.array([[0, 0, 1]], dtype=np.float32), (8, 1))
    F = np.array([[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7]], dtype=np.uint32)
    return Scene("stub_simple", P, N, F, meta={"uv": np.array([[0, 0], [1, 0], [1, 1], [0, 1]] * 2, dtype=np.float32)})


def stub_simple2_mesh() -> Scene:
    """StubTriangleMesh_Simple2 (stub.trianglemesh.h:124-181): the unit square on the plane x = -z."""
    P = np.array([[0, 0, 0], [1, 0, -1], [1, 1, -1], [0, 1, 0]], dtype=np.float32)
    n = np.array([1, 0, 1], dtype=np.float64) / np.sqrt(2.0)
    N = np.tile(n.astype(np.float32), (4, 1))
    F = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32)
    return Scene("stub_simple2", P, N, F, meta={"uv": np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.float32)})


def stub_random_mesh(values9000: np.ndarray) -> Scene:
    """StubTriangleMesh_Random (stub.trianglemesh.h:184-225) from its 9000 mt19937(42) draws."""
    P = np.asarray(values9000, dtype=np.float64).astype(np.float32).reshape(-1, 3)
    F = np.arange(P.shape[0], dtype=np.uint32).reshape(-1, 3)
    p1, p2, p3 = P[0::3], P[1::3], P[2::3]
    n = np.cross(p2 - p1, p3 - p1).astype(np.float32)
    N = np.repeat(n, 3, axis=0)
    return Scene("stub_random", P, N, F)


def grid_rays_simple() -> Rays:
    """The 81 rays of Intersect_Simple: from (0,0,1) toward (x, y, 0), x, y in {0.1..0.9} (test.scene.intersection.cpp:117-150)."""
    xs = (np.arange(1, 10, dtype=np.float32) / np.float32(10))
    o, d = [], []
    for y in xs:
        for x in xs:
            v = np.array([x, y, 0], dtype=np.float32) - np.array([0, 0, 1], dtype=np.float32)
            o.append([0, 0, 1])
            d.append(v / np.float32(np.sqrt(np.float32(v @ v))))
    n = len(o)
    return Rays(np.asarray(o, dtype=np.float32), np.asarray(d, dtype=np.float32), np.zeros(n, np.float32),
                np.full(n, FLT_MAX, np.float32))


def grid_rays_down() -> Rays:
    """The 81 rays of Intersect_Simple2 / Consistency: from (x, y, 1) along -z (test.scene.intersection.cpp:152-228)."""
    xs = (np.arange(1, 10, dtype=np.float32) / np.float32(10))
    o = np.array([[x, y, 1] for y in xs for x in xs], dtype=np.float32)
    d = np.tile(np.array([[0, 0, -1]], dtype=np.float32), (len(o), 1))
    n = len(o)
    return Rays(o, d, np.zeros(n, np.float32), np.full(n, FLT_MAX, np.float32))
