"""Seeded procedural scenes for the pt workloads, and their pbrt-v4 form.

Every scene is world-space geometry (instances baked), materials in the reference's pbrt
mapping onto OpenPBRMaterial (src/merian-shaders/scene/pbrt_scene.cpp:365-487):
  diffuse    -> base = reflectance, specular_weight 0, specular_alpha (1, 1)
  conductor  -> metalness 1, base = reflectance (F0), specular_alpha = roughness (remaproughness false)
  dielectric -> specular_ior = eta, transmission {1, white}, TwoSided
and environments as the pbrt loader builds them (constant infinite light -> 1x1 equal-area
octahedral texture, clamp-to-edge, pbrt_scene.cpp:924-969; --env-map -> lat-long, repeat).

Workloads (sizes as named):
  plain_small       8x8 icospheres (1,280 tris each) + ground, 512x512, 4 spp, L = 4
  ray_query         1M random triangles in the unit cube, 16M incoherent + 4096^2 coherent rays
  full_1080p        256 displaced tori instances (~2M tris) in an enclosing room lit through a window and a skylight, 256 materials,
                    2048x1024 lat-long sky + sun, 1920x1080, 64 spp, L = 8
  serving scenes    8 scenes of 100k..2M triangles built like the two above
"""
from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

F32 = np.float32


# ------------------------------------------------------------------------------------------
# encodings (reference include/merian-shaders/utils/encoding.slang:6-55), float32 op by op

def encode_normals(n: np.ndarray) -> np.ndarray:
    """packSnorm2x16(octahedral_project(n)) for an (N, 3) float32 array."""
    n = np.asarray(n, dtype=F32)
    x, y, z = n[:, 0], n[:, 1], n[:, 2]
    inv = F32(1.0) / ((np.abs(x) + np.abs(y)) + np.abs(z))
    px = np.where(z < 0, np.copysign(F32(1.0) - np.abs(y * inv), x), x * inv).astype(F32)
    py = np.where(z < 0, np.copysign(F32(1.0) - np.abs(x * inv), y), y * inv).astype(F32)
    qx = np.rint(np.clip(px, F32(-1), F32(1)) * F32(32767.0)).astype(np.int16)
    qy = np.rint(np.clip(py, F32(-1), F32(1)) * F32(32767.0)).astype(np.int16)
    return qx.view(np.uint16).astype(np.uint32) | (qy.view(np.uint16).astype(np.uint32) << 16)


def f16_bits(a) -> np.ndarray:
    return np.asarray(a, dtype=np.float16).view(np.uint16)


# ------------------------------------------------------------------------------------------
# scene container

@dataclass
class CameraSpec:
    eye: tuple
    target: tuple
    up: tuple
    vfov: float          # radians, vertical
    width: int
    height: int
    near: float = 0.01
    far: float = 1000.0
    jitter: int = abi.JITTER_R2  # pbrt-loaded cameras jitter with R2 (pbrt_scene.cpp:1006-1008)

    @property
    def aspect(self) -> float:
        return float(F32(self.width) / F32(self.height))


@dataclass
class Scene:
    name: str
    vertices: np.ndarray        # VERTEX_DTYPE
    indices: np.ndarray         # uint32, 3 per triangle, local to each geometry's vertices
    geometries: np.ndarray      # GEOMETRY_DTYPE
    materials: np.ndarray       # MATERIAL_DTYPE
    env_type: int = abi.ENV_NONE
    env_wrap: int = abi.WRAP_CLAMP
    env_rgba: np.ndarray | None = None  # (H, W, 4) float32
    env_intensity: float = 1.0
    env_to_local: np.ndarray = field(default_factory=lambda: np.eye(3, dtype=F32))
    camera: CameraSpec | None = None
    _keep: list = field(default_factory=list, repr=False)

    @property
    def triangle_count(self) -> int:
        return int(self.geometries["primitive_count"].sum())

    def bounds(self):
        """(min xyz, max xyz) of the vertex positions."""
        p = self.vertices["position"]
        return p.min(0), p.max(0)

    def desc(self) -> abi.SceneDesc:
        """mrt_scene_desc pointing into this scene's arrays (kept alive by the scene)."""
        d = abi.SceneDesc()
        v = np.ascontiguousarray(self.vertices)
        i = np.ascontiguousarray(self.indices, dtype=np.uint32)
        g = np.ascontiguousarray(self.geometries)
        m = np.ascontiguousarray(self.materials)
        self._keep = [v, i, g, m]
        d.vertices, d.vertex_count = abi.ptr(v), len(v)
        d.indices, d.index_count = abi.ptr(i), len(i)
        d.geometries, d.geometry_count = abi.ptr(g), len(g)
        d.materials, d.material_count = abi.ptr(m), len(m)
        d.env.type = self.env_type
        d.env.wrap = self.env_wrap
        if self.env_type != abi.ENV_NONE:
            e = np.ascontiguousarray(self.env_rgba, dtype=F32)
            self._keep.append(e)
            d.env.height, d.env.width = e.shape[0], e.shape[1]
            d.env.rgba = abi.ptr(e)
        d.env.intensity = self.env_intensity
        for k, val in enumerate(np.asarray(self.env_to_local, dtype=F32).reshape(9)):
            d.env.to_local[k] = float(val)
        return d

    def digest(self) -> str:
        h = hashlib.sha256()
        for a in (self.vertices, self.indices, self.geometries, self.materials):
            h.update(np.ascontiguousarray(a).tobytes())
        if self.env_rgba is not None:
            h.update(np.ascontiguousarray(self.env_rgba, dtype=F32).tobytes())
        return h.hexdigest()[:16]


class SceneBuilder:
    """Accumulates world-space meshes into the shared arrays of a Scene."""

    def __init__(self, name: str):
        self.name = name
        self.verts: list[np.ndarray] = []
        self.idx: list[np.ndarray] = []
        self.geoms: list[tuple] = []
        self.mats: list[np.ndarray] = []
        self._nv = 0
        self._ni = 0

    def add_material(self, m: np.ndarray) -> int:
        self.mats.append(m)
        return len(self.mats) - 1

    def add_mesh(self, pos, nrm, tri, material: int, flags: int = 0, mask: int = 0xFF):
        pos = np.asarray(pos, dtype=F32)
        tri = np.asarray(tri, dtype=np.uint32).reshape(-1, 3)
        v = np.zeros(len(pos), dtype=abi.VERTEX_DTYPE)
        v["position"] = pos
        v["encoded_normal"] = encode_normals(np.asarray(nrm, dtype=F32))
        self.verts.append(v)
        self.idx.append(tri.reshape(-1))
        self.geoms.append((self._nv, self._ni, len(tri), material,
                           flags | abi.GEOM_PRETRANSFORMED | abi.GEOM_INDEX_UINT32, mask))
        self._nv += len(pos)
        self._ni += tri.size

    def build(self, **kw) -> Scene:
        g = np.zeros(len(self.geoms), dtype=abi.GEOMETRY_DTYPE)
        for k, (vo, io, pc, mid, fl, mk) in enumerate(self.geoms):
            g[k] = (vo, io, pc, mid, fl, mk)
        return Scene(self.name, np.concatenate(self.verts), np.concatenate(self.idx).astype(np.uint32),
                     g, np.concatenate(self.mats) if self.mats else np.zeros(1, abi.MATERIAL_DTYPE), **kw)


# ------------------------------------------------------------------------------------------
# materials (pbrt mapping, pbrt_scene.cpp:365-487; OpenPBR defaults openpbr_material.hpp:54-62)

def _openpbr_defaults() -> np.ndarray:
    m = np.zeros(1, dtype=abi.MATERIAL_DTYPE)
    m["model"] = abi.MAT_OPENPBR
    m["base_color"] = (1, 1, 1)
    m["opacity"] = 1.0
    m["specular_alpha"] = (1.0, 1.0)
    m["specular_weight"] = 1.0
    m["specular_ior"] = 1.5
    m["coat_ior"] = 1.6
    m["sheen_color"] = (1, 1, 1)
    m["sheen_alpha"] = 0.09
    return m


def pbrt_diffuse(reflectance) -> np.ndarray:
    m = _openpbr_defaults()
    m["base_color"] = reflectance
    m["specular_weight"] = 0.0
    m["specular_alpha"] = (1.0, 1.0)
    return m


def pbrt_conductor(reflectance, alpha: float) -> np.ndarray:
    m = _openpbr_defaults()
    m["metalness"] = 1.0
    m["base_color"] = reflectance
    m["specular_alpha"] = (alpha, alpha)
    return m


def pbrt_dielectric(eta: float = 1.5, alpha: float = 0.0) -> np.ndarray:
    m = _openpbr_defaults()
    m["features"] = abi.PBR_TRANSMISSION
    m["specular_ior"] = eta
    m["transmission_weight"] = 1.0
    m["transmission_color"] = (1, 1, 1)
    m["specular_alpha"] = (alpha, alpha)
    return m


def diffuse_material(albedo) -> np.ndarray:
    """DiffuseMaterial (diffuse-material.slang:7-43)."""
    m = np.zeros(1, dtype=abi.MATERIAL_DTYPE)
    m["model"] = abi.MAT_DIFFUSE
    m["base_color"] = albedo
    m["opacity"] = 1.0
    return m


# ------------------------------------------------------------------------------------------
# meshes

def icosphere(subdiv: int):
    """Unit icosphere, outward winding (cross(e01, e02) points out), float64."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = f
    for _ in range(subdiv):
        cache = {}
        nf = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                p = verts[a] + verts[b]
                verts.append(p / np.linalg.norm(p))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    P = np.array(verts)
    T = np.array(faces, dtype=np.int64)
    # orient outward
    n = np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]])
    flip = (n * (P[T[:, 0]] + P[T[:, 1]] + P[T[:, 2]])).sum(1) < 0
    T[flip] = T[flip][:, [0, 2, 1]]
    return P, T


def torus(nu: int, nv: int, R: float, r: float):
    """Torus around +y with nu x nv quads (2*nu*nv triangles), outward winding, float64."""
    u = np.arange(nu) * (2 * np.pi / nu)
    v = np.arange(nv) * (2 * np.pi / nv)
    U, V = np.meshgrid(u, v, indexing="ij")
    P = np.stack([(R + r * np.cos(V)) * np.cos(U), r * np.sin(V), (R + r * np.cos(V)) * np.sin(U)], -1)
    P = P.reshape(-1, 3)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    a = (i * nv + j).reshape(-1)
    b = (((i + 1) % nu) * nv + j).reshape(-1)
    c = (((i + 1) % nu) * nv + (j + 1) % nv).reshape(-1)
    d = (i * nv + (j + 1) % nv).reshape(-1)
    T = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    n = np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]])
    centre_ring = np.stack([np.cos(np.arctan2(P[T[:, 0], 2], P[T[:, 0], 0])) * R, np.zeros(len(T)),
                            np.sin(np.arctan2(P[T[:, 0], 2], P[T[:, 0], 0])) * R], 1)
    out = P[T[:, 0]] - centre_ring
    flip = (n * out).sum(1) < 0
    T[flip] = T[flip][:, [0, 2, 1]]
    return P, T


def vertex_normals(P: np.ndarray, T: np.ndarray) -> np.ndarray:
    n = np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]])
    vn = np.zeros_like(P)
    for k in range(3):
        np.add.at(vn, T[:, k], n)
    vn /= np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-30)
    return vn


def quad(a, b, c, d):
    """Two triangles (a, b, c), (a, c, d); front = cross(b - a, c - a)."""
    P = np.array([a, b, c, d], dtype=np.float64)
    T = np.array([[0, 1, 2], [0, 2, 3]])
    n = np.cross(P[1] - P[0], P[2] - P[0])
    N = np.tile(n / np.linalg.norm(n), (4, 1))
    return P, N, T


def rotation(axis, angle):
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    x, y, z = axis
    c, s = math.cos(angle), math.sin(angle)
    return np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                     [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                     [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])


# ------------------------------------------------------------------------------------------
# environments

def constant_env(L=(1.0, 1.0, 1.0)) -> dict:
    """pbrt `LightSource "infinite" "rgb L"`: 1x1 equal-area octahedral texture, clamp."""
    rgba = np.array(list(L) + [1.0], dtype=F32).reshape(1, 1, 4)
    return dict(env_type=abi.ENV_EQUAL_AREA_OCT, env_wrap=abi.WRAP_CLAMP, env_rgba=rgba)


def sky_env(width=2048, height=1024, sun_dir=(0.35, 0.55, -0.76), dynamic_range=5e4) -> dict:
    """Lat-long sky: gradient + sun lobe, max/min radiance = dynamic_range (repeat wrap)."""
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    theta = v * np.pi                      # v = acos(y) / pi
    phi = (u - 0.5) * 2 * np.pi            # u = atan2(x, -z) / 2pi + 0.5
    TH, PH = np.meshgrid(theta, phi, indexing="ij")
    y = np.cos(TH)
    x = np.sin(TH) * np.sin(PH)
    z = -np.sin(TH) * np.cos(PH)
    elev = np.clip(y, 0, 1)[..., None]
    horizon = np.array([0.9, 0.9, 1.0])
    zenith = np.array([0.25, 0.45, 1.0])
    sky = horizon * (1 - elev) + zenith * elev
    ground = np.array([0.08, 0.07, 0.06])
    below = (y < 0)[..., None]
    base = np.where(below, ground * (1 + 0.5 * y[..., None]), sky)
    base_min = float(base.min())
    s = np.asarray(sun_dir, dtype=np.float64)
    s = s / np.linalg.norm(s)
    cosang = x * s[0] + y * s[1] + z * s[2]
    lobe = np.exp(2000.0 * (cosang - 1.0))  # ~1.5 degree sun
    sun_peak = dynamic_range * base_min - float(base[np.unravel_index(np.argmax(cosang), cosang.shape)].max())
    rgb = base + lobe[..., None] * sun_peak * np.array([1.0, 0.95, 0.85])
    rgba = np.concatenate([rgb, np.ones_like(rgb[..., :1])], -1).astype(F32)
    return dict(env_type=abi.ENV_LATLONG, env_wrap=abi.WRAP_REPEAT, env_rgba=rgba)


# ------------------------------------------------------------------------------------------
# workloads

def plain_small(seed: int = 0, res: int = 512, subdiv: int = 3, grid: int = 8) -> Scene:
    """8x8 icospheres (radius U[0.2, 0.5], pitch 1.25) on a 12x12 ground, Lambert / GGX metal."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder(f"plain_small_s{seed}")
    P0, T0 = icosphere(subdiv)
    N0 = vertex_normals(P0, T0)
    ground = b.add_material(pbrt_diffuse((0.5, 0.5, 0.5)))
    for gi in range(grid):
        for gj in range(grid):
            r = rng.uniform(0.2, 0.5)
            c = np.array([(gi - (grid - 1) / 2) * 1.25, r, (gj - (grid - 1) / 2) * 1.25])
            if rng.uniform() < 0.5:
                mid = b.add_material(pbrt_diffuse(rng.uniform(0.1, 0.9, 3)))
            else:
                refl = rng.uniform(0.1, 0.9, 3)
                rough = rng.uniform(0.05, 1.0)
                mid = b.add_material(pbrt_conductor(refl, rough * rough))
            b.add_mesh(P0 * r + c, N0, T0, mid)
    P, N, T = quad((-6, 0, -6), (-6, 0, 6), (6, 0, 6), (6, 0, -6))
    b.add_mesh(P, N, T, ground)
    cam = CameraSpec((0.0, 5.0, 10.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.radians(40.0), res, res)
    return b.build(camera=cam, **constant_env())


def ray_query(seed: int = 1, n_tri: int = 1_000_000):
    """Random triangles in the unit cube: centroid U[0,1]^3, size 10^U[-3,-1], vertices on a sphere."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 1, (n_tri, 1, 3))
    ell = 10.0 ** rng.uniform(-3, -1, (n_tri, 1, 1))
    u = rng.normal(size=(n_tri, 3, 3))
    u /= np.linalg.norm(u, axis=2, keepdims=True)
    u -= u.mean(axis=1, keepdims=True)
    pos = (c + ell * u).astype(F32).reshape(-1, 3)
    idx = np.arange(3 * n_tri, dtype=np.uint32)
    return pos, idx


def incoherent_rays(n: int, seed: int = 11) -> np.ndarray:
    rng = np.random.default_rng(seed)
    r = np.zeros(n, dtype=abi.RAY_DTYPE)
    r["origin"] = rng.uniform(0, 1, (n, 3)).astype(F32)
    z = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    s = np.sqrt(np.maximum(0.0, 1 - z * z))
    r["direction"] = np.stack([s * np.cos(phi), s * np.sin(phi), z], 1).astype(F32)
    r["tmin"] = 0.0
    r["tmax"] = np.inf
    return r


def ray_query_camera(res: int = 4096) -> CameraSpec:
    return CameraSpec((0.5, 0.5, -1.0), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), math.radians(60.0), res, res,
                      near=0.0, far=float("inf"), jitter=abi.JITTER_NONE)


def _displaced_torus(rng, nu=63, nv=62):
    P, T = torus(nu, nv, 1.0, 0.38)
    k = rng.integers(2, 6, 3)
    ph = rng.uniform(0, 2 * np.pi, 3)
    ang_u = np.arctan2(P[:, 2], P[:, 0])
    ang_v = np.arctan2(P[:, 1], np.hypot(P[:, 0], P[:, 2]) - 1.0)
    disp = 0.07 * np.sin(k[0] * ang_u + ph[0]) * np.cos(k[1] * ang_v + ph[1]) + \
        0.04 * np.sin(k[2] * (ang_u + ang_v) + ph[2])
    ring = np.stack([np.cos(ang_u), np.zeros_like(ang_u), np.sin(ang_u)], 1)
    radial = P - ring
    radial /= np.linalg.norm(radial, axis=1, keepdims=True)
    P = P + disp[:, None] * radial
    return P, T, vertex_normals(P, T)


def _rect_with_hole(o, U, V, hole):
    """Quads covering the rectangle o + s U + t V, s, t in [0, 1], minus the hole
    (s0, s1, t0, t1) if given; fronts along cross(U, V)."""
    def rect(s0, s1, t0, t1):
        return quad(o + s0 * U + t0 * V, o + s1 * U + t0 * V, o + s1 * U + t1 * V, o + s0 * U + t1 * V)
    if hole is None:
        return [rect(0, 1, 0, 1)]
    s0, s1, t0, t1 = hole
    return [rect(0, 1, 0, t0), rect(0, 1, t1, 1), rect(0, s0, t0, t1), rect(s1, 1, t0, t1)]


def full_scene(seed: int = 2, instances: int = 256, res=(1920, 1080), env_size=(2048, 1024),
               nu: int = 63, nv: int = 62, name: str = "full_1080p", enclosed: bool = True,
               dynamic_range: float = 5e4) -> Scene:
    """Displaced tori instances in an enclosing room lit through a window and a skylight, 256
    materials, sky env with a sun lobe.

    enclosed=False gives the open set instead (floor and three walls under the open sky, camera
    outside): the sun reaches most surfaces directly, which tests of estimators that rely on
    sampling the sun by the BSDF alone need. dynamic_range is the sky's max/min radiance (a
    weaker sun for tests whose estimators would otherwise be heavy-tailed)."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder(f"{name}_s{seed}")
    bases = [_displaced_torus(rng, nu, nv) for _ in range(4)]
    mats = []
    for k in range(instances):
        r = rng.uniform()
        if r < 0.4:
            mats.append(pbrt_diffuse(rng.uniform(0.1, 0.9, 3)))
        elif r < 0.75:
            rough = rng.uniform(0.05, 1.0)
            mats.append(pbrt_conductor(rng.uniform(0.2, 0.95, 3), rough * rough))
        else:
            rough = rng.uniform(0.0, 0.4)
            mats.append(pbrt_dielectric(1.5, rough * rough))
    mat_ids = [b.add_material(m) for m in mats]
    side = int(math.ceil(math.sqrt(instances)))
    for k in range(instances):
        P, T, N = bases[k % 4]
        s = rng.uniform(0.35, 0.6)
        R = rotation(rng.normal(size=3), rng.uniform(0, 2 * np.pi))
        gx, gz = k % side, k // side
        t = np.array([(gx - (side - 1) / 2) * 1.6 + rng.uniform(-0.3, 0.3), 0.9 + rng.uniform(0, 1.2),
                      (gz - (side - 1) / 2) * 1.6 + rng.uniform(-0.3, 0.3)])
        flags = abi.GEOM_TWO_SIDED if mats[k]["features"][0] & abi.PBR_TRANSMISSION else 0
        b.add_mesh((P * s) @ R.T + t, N @ R.T, T, mat_ids[k], flags)
    half = side * 0.8 + 2.0
    H = 6.0
    floor_m = b.add_material(pbrt_diffuse((0.6, 0.6, 0.6)))
    wall_m = b.add_material(pbrt_diffuse((0.7, 0.55, 0.45)))
    if not enclosed:
        P, N, T = quad((-half, 0, -half), (-half, 0, half), (half, 0, half), (half, 0, -half))
        b.add_mesh(P, N, T, floor_m)
        for q in (((-half, 0, -half), (half, 0, -half), (half, H, -half), (-half, H, -half)),   # back
                  ((-half, 0, half), (-half, 0, -half), (-half, H, -half), (-half, H, half)),   # left
                  ((half, 0, -half), (half, 0, half), (half, H, half), (half, H, -half))):      # right
            P, N, T = quad(*q)
            b.add_mesh(P, N, T, wall_m)
        cam = CameraSpec((0.0, 7.0, half + 9.0), (0.0, 0.8, 0.0), (0.0, 1.0, 0.0), math.radians(45.0),
                         res[0], res[1])
        return b.build(camera=cam, **sky_env(env_size[0], env_size[1], dynamic_range=dynamic_range))
    # the six faces of the room, fronts facing inwards (walls are one-sided): rect(o, U, V) spans
    # o + s U + t V with cross(U, V) as its front; the back wall (towards the sun) has a window and
    # the ceiling a skylight, the only ways light gets in
    X, Y, Z = np.array([2 * half, 0, 0]), np.array([0, H, 0]), np.array([0, 0, 2 * half])
    faces = [((-half, 0, -half), Z, X, None),                          # floor, front +y
             ((-half, H, -half), X, Z, (0.32, 0.68, 0.2, 0.55)),         # ceiling, -y, skylight
             ((-half, 0, -half), X, Y, (0.22, 0.78, 0.2, 0.8)),          # back wall z = -half, +z, window
             ((-half, 0, half), Y, X, None),                             # front wall, -z
             ((-half, 0, half), -Z, Y, None),                            # left wall x = -half, +x
             ((half, 0, -half), Z, Y, None)]                             # right wall, -x
    for k, (o, U, V, hole) in enumerate(faces):
        for P, N, T in _rect_with_hole(np.array(o, np.float64), U, V, hole):
            b.add_mesh(P, N, T, floor_m if k == 0 else wall_m)
    cam = CameraSpec((0.0, 4.0, half - 0.8), (0.0, 0.8, -0.25 * half), (0.0, 1.0, 0.0), math.radians(60.0),
                     res[0], res[1])
    return b.build(camera=cam, **sky_env(env_size[0], env_size[1], dynamic_range=dynamic_range))


def serving_scenes(count: int = 8):
    """Resident scenes of the serving mix (seeds 101..108, 103,682 to 1,999,896 triangles): icosphere
    grids (as the small plain-pt scene, 9x9) and tori in enclosing rooms (as the 1080p scene)."""
    scenes = []
    sizes = [13, 26, 52, 100, 160, 200, 230, 256]  # torus instances -> ~100k .. 2M triangles
    for k in range(count):
        seed = 101 + k
        if k % 4 == 0:
            scenes.append(plain_small(seed=seed, subdiv=3 + (k // 4), grid=9))
        else:
            scenes.append(full_scene(seed=seed, instances=sizes[k], env_size=(512, 256), name="serve"))
    return scenes


# ------------------------------------------------------------------------------------------
# pbrt-v4 writer: the same scene for the reference's "PBRT Scene" node (examples/pbrt.json)

def write_pbrt(scene: Scene, path: str, env_pfm: str | None = None):
    """Writes `scene` as pbrt-v4 text: world-space trianglemeshes with explicit normals."""
    cam = scene.camera
    lines = [f'LookAt {cam.eye[0]} {cam.eye[1]} {cam.eye[2]}  {cam.target[0]} {cam.target[1]} {cam.target[2]}  '
             f'{cam.up[0]} {cam.up[1]} {cam.up[2]}',
             f'Camera "perspective" "float fov" [{math.degrees(cam.vfov if cam.width >= cam.height else cam.vfov)}]',
             f'Film "rgb" "integer xresolution" [{cam.width}] "integer yresolution" [{cam.height}]',
             "WorldBegin"]
    if scene.env_type == abi.ENV_EQUAL_AREA_OCT and scene.env_rgba is not None and scene.env_rgba.shape[:2] == (1, 1):
        L = scene.env_rgba[0, 0, :3]
        lines.append(f'LightSource "infinite" "rgb L" [{L[0]} {L[1]} {L[2]}]')
    m = scene.materials
    for gi, g in enumerate(scene.geometries):
        mm = m[g["material_id"]]
        lines.append("AttributeBegin")
        if mm["features"] & abi.PBR_TRANSMISSION:
            lines.append(f'Material "dielectric" "float eta" [{mm["specular_ior"]}] '
                         f'"float roughness" [{mm["specular_alpha"][0]}] "bool remaproughness" false')
        elif mm["metalness"] > 0:
            c = mm["base_color"]
            lines.append(f'Material "conductor" "rgb reflectance" [{c[0]} {c[1]} {c[2]}] '
                         f'"float roughness" [{mm["specular_alpha"][0]}] "bool remaproughness" false')
        else:
            c = mm["base_color"]
            lines.append(f'Material "diffuse" "rgb reflectance" [{c[0]} {c[1]} {c[2]}]')
        vo, io, pc = int(g["vertex_offset"]), int(g["index_offset"]), int(g["primitive_count"])
        idx = scene.indices[io:io + 3 * pc]
        nv = int(idx.max()) + 1
        P = scene.vertices["position"][vo:vo + nv]
        lines.append('Shape "trianglemesh" "integer indices" [' + " ".join(map(str, idx.tolist())) + "] "
                     '"point3 P" [' + " ".join(f"{x:.9g}" for x in P.reshape(-1)) + "]")
        lines.append("AttributeEnd")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
