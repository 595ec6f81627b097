"""Scene descriptors for the C ABI, built from Python, and the seeded synthetic
scenes the benchmarks and tests use.

``Scene`` mirrors what Atrc's factory creates from JSON (entities with a
geometry, a material and an optional emission; an environment light; an
aggregate) and produces an ``agx_scene_desc`` that both the GPU library and
the CPU oracle accept.  Meshes are float32 arrays of shape (3*n, 8):
position, normal and uv of each triangle corner, already in world space (Atrc
transforms triangle_bvh meshes to world space at construction,
src/tracer/src/core/geometry/triangle_bvh.cpp:533-558).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import capi

IDENTITY = (1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0)


@dataclass
class Material:
    kind: int = capi.MAT_IDEAL_DIFFUSE
    color: tuple = (0.5, 0.5, 0.5)
    color2: tuple = (0.0, 0.0, 0.0)
    eta: tuple = (1.0, 1.0, 1.0)
    k: tuple = (0.0, 0.0, 0.0)
    roughness: float = 0.5
    anisotropic: float = 0.0
    ior: float = 1.5
    separate_refraction: bool = False
    # disney: color = base_color, color2 = specular_scale
    metallic: float = 0.0
    transmission: float = 0.0
    transmission_roughness: float | None = None   # None = roughness (material_creators.cpp:71-73)
    specular_tint: float = 0.0
    sheen: float = 0.0
    sheen_tint: float = 0.0
    clearcoat: float = 0.0
    clearcoat_gloss: float = 1.0


@dataclass
class Texture3D:
    """constant3d (value) or image3d (grid of shape (depth, height, width[, 3]), float32)."""
    value: tuple = (1.0, 1.0, 1.0)
    grid: np.ndarray | None = None
    linear: bool = True
    transform: tuple = IDENTITY          # uvw -> texture space, row-major 3x4
    wrap: tuple = (capi.WRAP_CLAMP, capi.WRAP_CLAMP, capi.WRAP_CLAMP)
    inv_gamma: float = 1.0


@dataclass
class Medium:
    """homogeneous (sigma_a, sigma_s, g) or heterogeneous (density/albedo/g grids in the
    unit cube of local_to_world) participating medium."""
    kind: int = capi.MEDIUM_HOMOGENEOUS
    sigma_a: tuple = (0.0, 0.0, 0.0)
    sigma_s: tuple = (0.0, 0.0, 0.0)
    g: float = 0.0
    max_scattering_count: int = 2 ** 31 - 1
    local_to_world: tuple = IDENTITY
    density: Texture3D = field(default_factory=lambda: Texture3D(value=(1.0, 1.0, 1.0)))
    albedo: Texture3D = field(default_factory=lambda: Texture3D(value=(1.0, 1.0, 1.0)))
    g3: Texture3D = field(default_factory=lambda: Texture3D(value=(0.0, 0.0, 0.0)))
    white_for_indirect: bool = False


def _affine_pair(l2w):
    m = np.array(l2w, np.float64).reshape(3, 4)
    full = np.eye(4)
    full[:3, :] = m
    inv = np.linalg.inv(full)[:3, :]
    return ([float(x) for x in m.reshape(-1).astype(np.float32)],
            [float(x) for x in inv.reshape(-1).astype(np.float32)])


def _fill_tex3d(d, t: Texture3D, keep: list) -> None:
    d.linear = 1 if t.linear else 0
    d.value[:] = [float(x) for x in t.value]
    d.transform[:] = [float(x) for x in t.transform]
    d.wrap[:] = [int(x) for x in t.wrap]
    d.inv_gamma = float(t.inv_gamma)
    if t.grid is None:
        d.kind = capi.TEX3D_CONSTANT
        d.width = d.height = d.depth = d.channels = 1
        return
    g = np.ascontiguousarray(t.grid, np.float32)
    if g.ndim == 3:
        g = g[..., None]
    d.kind = capi.TEX3D_GRID
    d.depth, d.height, d.width, d.channels = (int(x) for x in g.shape)
    keep.append(g)
    d.data = g.ctypes.data_as(C.POINTER(C.c_float))


@dataclass
class Entity:
    geometry: int
    material: int
    emit: tuple = (0.0, 0.0, 0.0)
    tris: np.ndarray | None = None       # (3n, 8) float32 world-space mesh
    points: tuple = ()                   # quad A..D (world) / triangle A..C (local)
    uvs: tuple = ()
    radius: float = 1.0
    local_to_world: tuple = IDENTITY     # row-major 3x4
    no_denoise: bool = False
    double_sided: bool = False           # Atrc "double_sided" around the geometry
    med_in: int = -1                     # index into Scene.media, -1 = void
    med_out: int = -1


@dataclass
class Scene:
    entities: list = field(default_factory=list)
    materials: list = field(default_factory=list)
    # native_sky {"top": (r,g,b), "bottom": (r,g,b)}, or ibl {"image": (H, W, 3) float32
    # | "texel": (r,g,b) | "checker": (grid_count, color1, color2), optional "linear",
    # "wrap": (u, v), "transform": 2x3 row-major, "inv_gamma", "no_importance_sampling"}
    env: dict | None = None
    aggregate: int = capi.AGGREGATE_NATIVE
    max_leaf_size: int = 5
    media: list = field(default_factory=list)
    accel: int = capi.ACCEL_ATRC         # acceleration structure the GPU builds at upload

    def add_medium(self, m: Medium) -> int:
        self.media.append(m)
        return len(self.media) - 1

    def add_material(self, m: Material) -> int:
        self.materials.append(m)
        return len(self.materials) - 1

    def add(self, e: Entity) -> int:
        self.entities.append(e)
        return len(self.entities) - 1

    def triangle_count(self) -> int:
        return sum(len(e.tris) // 3 for e in self.entities if e.tris is not None)

    def to_desc(self):
        """Returns (agx_scene_desc, keepalive); keep both alive while the descriptor is used."""
        keep = []
        ents = (capi.agx_entity_desc * max(1, len(self.entities)))()
        for i, e in enumerate(self.entities):
            d = ents[i]
            d.geometry = e.geometry
            d.material = e.material
            d.medium_in = int(e.med_in)
            d.medium_out = int(e.med_out)
            d.emit_radiance[:] = [float(x) for x in e.emit]
            d.power = -1.0
            d.no_denoise = 1 if e.no_denoise else 0
            d.double_sided = 1 if e.double_sided else 0
            if e.tris is not None:
                arr = np.ascontiguousarray(e.tris, np.float32)
                keep.append(arr)
                d.tri_count = len(arr) // 3
                d.tris = arr.ctypes.data_as(C.POINTER(capi.agx_vertex))
            for k, p in enumerate(e.points):
                d.p[k][:] = [float(x) for x in p]
            for k, t in enumerate(e.uvs):
                d.t[k][:] = [float(x) for x in t]
            d.radius = float(e.radius)
            d.local_to_world[:], d.world_to_local[:] = _affine_pair(e.local_to_world)
        mats = (capi.agx_material_desc * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            d = mats[i]
            d.kind = m.kind
            d.color[:] = [float(x) for x in m.color]
            d.color2[:] = [float(x) for x in m.color2]
            d.eta[:] = [float(x) for x in m.eta]
            d.k[:] = [float(x) for x in m.k]
            d.roughness = m.roughness
            d.anisotropic = m.anisotropic
            d.ior = m.ior
            d.glass_separate_refraction = 1 if m.separate_refraction else 0
            d.metallic = m.metallic
            d.transmission = m.transmission
            d.transmission_roughness = m.roughness if m.transmission_roughness is None else m.transmission_roughness
            d.specular_tint = m.specular_tint
            d.sheen = m.sheen
            d.sheen_tint = m.sheen_tint
            d.clearcoat = m.clearcoat
            d.clearcoat_gloss = m.clearcoat_gloss
        desc = capi.agx_scene_desc()
        desc.abi_version = capi.AGX_ABI_VERSION
        desc.entity_count = len(self.entities)
        desc.entities = C.cast(ents, C.POINTER(capi.agx_entity_desc))
        desc.material_count = len(self.materials)
        desc.materials = C.cast(mats, C.POINTER(capi.agx_material_desc))
        if self.env is not None and "top" in self.env:
            desc.env.kind = capi.ENV_NATIVE_SKY
            desc.env.top[:] = [float(x) for x in self.env["top"]]
            desc.env.bottom[:] = [float(x) for x in self.env["bottom"]]
            desc.env.power = -1.0
        elif self.env is not None:
            e, t = self.env, desc.env.tex
            desc.env.kind = capi.ENV_IBL
            desc.env.power = -1.0
            desc.env.no_importance_sampling = 1 if e.get("no_importance_sampling") else 0
            if "image" in e:
                img = np.ascontiguousarray(e["image"], np.float32)
                keep.append(img)
                t.kind = capi.TEX2D_IMAGE
                t.height, t.width = img.shape[0], img.shape[1]
                t.data = img.ctypes.data_as(C.POINTER(C.c_float))
            elif "checker" in e:
                n, c1, c2 = e["checker"]
                t.kind, t.grid_count = capi.TEX2D_CHECKER, int(n)
                t.value[:] = [float(x) for x in c1]
                t.value2[:] = [float(x) for x in c2]
            else:
                t.kind = capi.TEX2D_CONSTANT
                t.value[:] = [float(x) for x in e["texel"]]
            t.linear = 0 if e.get("linear") is False else 1
            wrap = {"clamp": capi.WRAP_CLAMP, "repeat": capi.WRAP_REPEAT, "mirror": capi.WRAP_MIRROR}
            t.wrap[:] = [wrap[w] for w in e.get("wrap", ("clamp", "clamp"))]
            t.transform[:] = [float(x) for x in e.get("transform", (1, 0, 0, 0, 1, 0))]
            t.inv_gamma = float(e.get("inv_gamma", 1.0))
        desc.aggregate = self.aggregate
        desc.max_leaf_size = self.max_leaf_size
        meds = (capi.agx_medium_desc * max(1, len(self.media)))()
        for i, m in enumerate(self.media):
            d = meds[i]
            d.kind = m.kind
            d.max_scattering_count = int(m.max_scattering_count)
            d.sigma_a[:] = [float(x) for x in m.sigma_a]
            d.sigma_s[:] = [float(x) for x in m.sigma_s]
            d.g = float(m.g)
            d.local_to_world[:], d.world_to_local[:] = _affine_pair(m.local_to_world)
            _fill_tex3d(d.density, m.density, keep)
            _fill_tex3d(d.albedo, m.albedo, keep)
            _fill_tex3d(d.g3, m.g3, keep)
            d.white_for_indirect = 1 if m.white_for_indirect else 0
        desc.accel = int(self.accel)
        desc.medium_count = len(self.media)
        desc.media = C.cast(meds, C.POINTER(capi.agx_medium_desc))
        keep += [ents, mats, meds]
        return desc, keep


def render_params(width: int, height: int, spp: int, camera: dict, *, min_depth: int = 5, max_depth: int = 10,
                  cont_prob: float = 0.9, specular_depth: int = 20, eps: float = 3e-4, seed: int = 42,
                  filter_radius: float = 0.5, collect_stats: bool = False, sample_begin: int = 0,
                  guide: bool = False, guide_alpha: float = 0.5, guide_train: bool = True, guide_batch: int = 65536,
                  guide_steps: int = 4, guide_lr: float = 1e-2, guide_reset: bool = False, guide_seed: int = 0,
                  filter: str = "box", filter_alpha: float = 0.0, task_grid_size: int = 32, use_mis: bool = True,
                  guide_medium_depth: int = -1):
    p = capi.agx_render_params()
    p.width, p.height, p.spp = width, height, spp
    p.min_depth, p.max_depth, p.specular_depth = min_depth, max_depth, specular_depth
    p.cont_prob = cont_prob
    p.use_mis = 1 if use_mis else 0
    p.eps = eps
    p.filter_radius = filter_radius
    p.seed = seed
    p.camera.pos[:] = camera["pos"]
    p.camera.dst[:] = camera["dst"]
    p.camera.up[:] = camera.get("up", (0, 0, 1))
    p.camera.fov_rad = math.radians(camera.get("fov", 40.0))
    p.camera.lens_radius = camera.get("lens_radius", 0.0)
    p.camera.focal_distance = camera.get("focal_distance", 1.0)
    p.camera.film_aspect = width / height
    p.collect_stats = 1 if collect_stats else 0
    p.sample_begin = sample_begin
    p.guide = 1 if guide else 0
    p.guide_alpha = guide_alpha
    p.guide_train = 1 if guide_train else 0
    p.guide_batch = guide_batch
    p.guide_steps = guide_steps
    p.guide_lr = guide_lr
    p.guide_reset = 1 if guide_reset else 0
    p.guide_seed = guide_seed
    p.filter = {"box": capi.FILTER_BOX, "gaussian": capi.FILTER_GAUSSIAN}[filter]
    p.filter_alpha = filter_alpha
    p.task_grid_size = task_grid_size
    p.guide_medium_depth = guide_medium_depth
    return p


# ---------------------------------------------------------------------------------- meshes

def mesh_from_indexed(pos: np.ndarray, nor: np.ndarray, faces: np.ndarray) -> np.ndarray:
    """(3n, 8) vertex records from indexed positions/normals (uv = 0)."""
    out = np.zeros((len(faces) * 3, 8), np.float32)
    idx = faces.reshape(-1)
    out[:, 0:3] = pos[idx]
    out[:, 3:6] = nor[idx]
    return out


def box_mesh(lo, hi, rot_deg: float = 0.0) -> np.ndarray:
    """Axis box [lo, hi] rotated about its vertical (z) centre axis: 12 triangles,
    flat outward normals."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    c = (lo + hi) / 2
    corners = np.array([[x, y, z] for z in (lo[2], hi[2]) for y in (lo[1], hi[1]) for x in (lo[0], hi[0])])
    a = math.radians(rot_deg)
    R = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])
    corners = (corners - c) @ R.T + c
    quads = [(0, 2, 3, 1, (0, 0, -1)), (4, 5, 7, 6, (0, 0, 1)), (0, 1, 5, 4, (0, -1, 0)),
             (2, 6, 7, 3, (0, 1, 0)), (0, 4, 6, 2, (-1, 0, 0)), (1, 3, 7, 5, (1, 0, 0))]
    tris = []
    for q0, q1, q2, q3, n in quads:
        nn = R @ np.array(n, np.float64)
        for (i, j, k) in ((q0, q1, q2), (q0, q2, q3)):
            for v in (i, j, k):
                tris.append([*corners[v], *nn, 0.0, 0.0])
    return np.array(tris, np.float32)


def icosphere(subdiv: int):
    """Unit icosphere: (vertices, faces); 20 * 4^subdiv faces."""
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = f
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts), np.array(faces, np.int64)


def _value_noise(p: np.ndarray, seed: int) -> np.ndarray:
    """Seeded trilinear lattice value noise in [-1, 1]."""
    pi = np.floor(p).astype(np.int64)
    f = p - pi
    w = f * f * (3 - 2 * f)

    def h(ix, iy, iz):
        x = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761)
        x = (x ^ (x >> 13)) * 1274126177
        x = x ^ (x >> 16)
        return (x & 0xFFFF).astype(np.float64) / 32767.5 - 1.0

    out = 0.0
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                wx = w[:, 0] if dx else 1 - w[:, 0]
                wy = w[:, 1] if dy else 1 - w[:, 1]
                wz = w[:, 2] if dz else 1 - w[:, 2]
                out = out + wx * wy * wz * h(pi[:, 0] + dx, pi[:, 1] + dy, pi[:, 2] + dz)
    return out


def vertex_normals(pos: np.ndarray, faces: np.ndarray) -> np.ndarray:
    fn = np.cross(pos[faces[:, 1]] - pos[faces[:, 0]], pos[faces[:, 2]] - pos[faces[:, 0]])
    n = np.zeros_like(pos)
    for k in range(3):
        np.add.at(n, faces[:, k], fn)
    return n / np.linalg.norm(n, axis=1, keepdims=True)


# ---------------------------------------------------------------------------------- scenes

def cornell_box() -> tuple[Scene, dict, dict]:
    """Cornell-box-like scene: 5 wall quads, 2 rotated 12-triangle boxes, a ceiling
    quad light of radiance (17, 12, 4); z is up.  Returns (scene, camera, trace settings)."""
    L = 5.55
    s = Scene(aggregate=capi.AGGREGATE_BVH)
    white = s.add_material(Material(color=(0.75, 0.75, 0.75)))
    red = s.add_material(Material(color=(0.63, 0.06, 0.05)))
    green = s.add_material(Material(color=(0.14, 0.45, 0.09)))
    quad = lambda a, b, c, d, m, emit=(0, 0, 0): s.add(Entity(capi.GEOM_QUAD, m, emit=emit, points=(a, b, c, d)))
    quad((0, 0, 0), (L, 0, 0), (L, L, 0), (0, L, 0), white)          # floor, normal +z
    quad((0, 0, L), (0, L, L), (L, L, L), (L, 0, L), white)          # ceiling, -z
    quad((0, L, 0), (L, L, 0), (L, L, L), (0, L, L), white)          # back wall, -y
    quad((0, 0, 0), (0, L, 0), (0, L, L), (0, 0, L), red)            # left wall, +x
    quad((L, 0, 0), (L, 0, L), (L, L, L), (L, L, 0), green)          # right wall, -x
    s.add(Entity(capi.GEOM_TRIANGLE_BVH, white, tris=box_mesh((1.0, 0.9, 0.0), (2.65, 2.55, 1.65), 15.0)))
    s.add(Entity(capi.GEOM_TRIANGLE_BVH, white, tris=box_mesh((2.9, 2.7, 0.0), (4.55, 4.35, 3.3), -18.0)))
    h = L - 0.01
    quad((2.125, 2.25, h), (2.125, 3.3, h), (3.425, 3.3, h), (3.425, 2.25, h), white, emit=(17, 12, 4))
    camera = {"pos": (L / 2, -8.0, L / 2), "dst": (L / 2, 0.0, L / 2), "up": (0, 0, 1), "fov": 40.0}
    trace = {"min_depth": 2, "max_depth": 5, "cont_prob": 0.9}
    return s, camera, trace


def cornell_delta(disney: bool = False) -> tuple[Scene, dict, dict]:
    """The Cornell-like scene with delta surfaces and curved emitters, seen
    through a thin lens: a glass sphere (ior 1.5, tinted refraction), a
    mirror quad on the back wall, a small sphere light and a disk light
    beside the ceiling quad light; the camera has lens radius 0.08 focused
    on the glass sphere.  disney=True adds a Disney sphere (metallic +
    clearcoat), so the full shading kernel takes the delta branches too."""
    s, cam, tr = cornell_box()
    glass = s.add_material(Material(kind=capi.MAT_GLASS, color=(1.0, 1.0, 1.0), color2=(0.8, 0.9, 1.0), ior=1.5))
    mirror = s.add_material(Material(kind=capi.MAT_MIRROR, color=(0.9, 0.9, 0.9), eta=(0.2, 0.2, 0.2),
                                     k=(3.0, 3.0, 3.0)))
    black = s.add_material(Material(kind=capi.MAT_IDEAL_BLACK))
    s.add(Entity(capi.GEOM_SPHERE, glass, radius=0.7, local_to_world=(1, 0, 0, 1.6, 0, 1, 0, 1.4, 0, 0, 1, 2.4)))
    y = 5.55 - 0.02
    s.add(Entity(capi.GEOM_QUAD, mirror, points=((3.2, y, 0.6), (4.8, y, 0.6), (4.8, y, 2.6), (3.2, y, 2.6))))
    s.add(Entity(capi.GEOM_SPHERE, black, emit=(6.0, 8.0, 10.0), radius=0.18,
                 local_to_world=(1, 0, 0, 0.9, 0, 1, 0, 4.2, 0, 0, 1, 4.6)))
    # disk facing down (local +z turned to -z) below the ceiling near the right wall
    s.add(Entity(capi.GEOM_DISK, black, emit=(3.0, 6.0, 3.0), radius=0.35,
                 local_to_world=(1, 0, 0, 4.6, 0, -1, 0, 1.2, 0, 0, -1, 5.3)))
    if disney:
        d = s.add_material(Material(kind=capi.MAT_DISNEY, color=(0.9, 0.6, 0.3), metallic=0.7, roughness=0.3,
                                    clearcoat=0.8, clearcoat_gloss=0.7))
        s.add(Entity(capi.GEOM_SPHERE, d, radius=0.5, local_to_world=(1, 0, 0, 4.0, 0, 1, 0, 1.4, 0, 0, 1, 0.5)))
    cam = dict(cam, lens_radius=0.08, focal_distance=9.0)
    return s, cam, dict(tr, max_depth=6, specular_depth=8)


def _lattice_hash(ix, iy, iz, seed: int) -> np.ndarray:
    """Value of the noise lattice at integer points (in [-1, 1])."""
    x = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761)
    x = (x ^ (x >> 13)) * 1274126177
    x = x ^ (x >> 16)
    return (x & 0xFFFF).astype(np.float64) / 32767.5 - 1.0


def _value_noise_grid(c: np.ndarray, seed: int) -> np.ndarray:
    """_value_noise over the tensor grid c x c x c (indexed [z, y, x]): the same
    trilinear lattice noise, evaluated one axis at a time."""
    pi = np.floor(c).astype(np.int64)
    f = c - pi
    w = f * f * (3 - 2 * f)
    lo = int(pi.min())
    k = np.arange(lo, int(pi.max()) + 2)
    L = _lattice_hash(k[:, None, None], k[None, :, None], k[None, None, :], seed)   # [ix, iy, iz]
    idx = pi - lo
    wts = np.stack([1 - w, w], axis=1)                                              # (n, 2)

    def along(a, axis):
        a = np.moveaxis(a, axis, -1)
        out = a[..., idx] * wts[:, 0] + a[..., idx + 1] * wts[:, 1]
        return np.moveaxis(out, -1, axis)

    g = along(along(along(L, 0), 1), 2)      # [x, y, z]
    return np.ascontiguousarray(g.transpose(2, 1, 0))


def fbm_grid(n: int, seed: int, octaves: int = 4, base: float = 4.0) -> np.ndarray:
    """Seeded n^3 fBm density in [0, 1] (value noise, 'octaves' octaves, lowest frequency
    'base' cells across the cube), indexed [z, y, x]."""
    c = (np.arange(n, dtype=np.float64) + 0.5) / n
    acc = np.zeros((n, n, n))
    amp, freq, norm = 1.0, base, 0.0
    for o in range(octaves):
        acc += amp * _value_noise_grid(c * freq, seed + 17 * o)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return np.clip(0.5 + 0.5 * acc / norm * 1.6, 0.0, 1.0).astype(np.float32)


def _box_frame(lo, hi, rot_deg: float) -> tuple:
    """3x4 local_to_world taking the unit cube onto box_mesh(lo, hi, rot_deg)."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    a = math.radians(rot_deg)
    R = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])
    Lm = R @ np.diag(hi - lo)
    t = (lo + hi) / 2 - Lm @ np.array([0.5, 0.5, 0.5])
    return tuple(float(v) for v in np.concatenate([Lm, t[:, None]], axis=1).reshape(-1))


def volumetric_cornell(grid_n: int = 16, seed: int = 5, sky: bool = True) -> tuple[Scene, dict, dict]:
    """The Cornell-like scene with both boxes turned into medium boundaries
    (invisible_surface): the short box holds a homogeneous medium (chromatic
    sigma_a/sigma_s, HG g = -0.2), the tall one a heterogeneous medium on a
    seeded grid_n^3 fBm density (max 2), constant albedo 0.9 and g = 0.3; a dim
    native_sky lights the room through its open front."""
    s, cam, tr = cornell_box()
    inv = s.add_material(Material(kind=capi.MAT_INVISIBLE_SURFACE))
    homo = s.add_medium(Medium(kind=capi.MEDIUM_HOMOGENEOUS, sigma_a=(0.05, 0.1, 0.2), sigma_s=(0.6, 0.5, 0.4),
                               g=-0.2))
    dens = fbm_grid(grid_n, seed) * 2.0
    het = s.add_medium(Medium(kind=capi.MEDIUM_HETEROGENEOUS,
                              local_to_world=_box_frame((2.9, 2.7, 0.0), (4.55, 4.35, 3.3), -18.0),
                              density=Texture3D(grid=dens),
                              albedo=Texture3D(value=(0.9, 0.9, 0.9)),
                              g3=Texture3D(value=(0.3, 0.3, 0.3))))
    short, tall = s.entities[5], s.entities[6]
    short.material, short.med_in = inv, homo
    tall.material, tall.med_in = inv, het
    if sky:
        s.env = {"top": (0.3, 0.35, 0.45), "bottom": (0.1, 0.1, 0.1)}
    return s, cam, dict(tr, max_depth=10)


def medium_furnace(albedo_one: bool = True, n_tris_subdiv: int = 2, radius: float = 2.0,
                   sigma_s: float = 0.8, heterogeneous: bool = False) -> tuple[Scene, dict, dict]:
    """Closed emitter: an inward-facing icosphere of radius `radius` emitting radiance 1
    (ideal_black) filled with a medium of albedo 1 (sigma_a = 0).  The radiance field
    inside is exactly 1 everywhere; the camera sits at the centre."""
    s = Scene(aggregate=capi.AGGREGATE_NATIVE)
    black = s.add_material(Material(kind=capi.MAT_IDEAL_BLACK))
    pos, faces = icosphere(n_tris_subdiv)
    pos = pos * radius
    faces = faces[:, ::-1].copy()                    # inward winding: geometric normal points to the centre
    tris = mesh_from_indexed(pos, -pos / radius, faces)
    if heterogeneous:
        m = Medium(kind=capi.MEDIUM_HETEROGENEOUS,
                   local_to_world=(2 * radius, 0, 0, -radius, 0, 2 * radius, 0, -radius, 0, 0, 2 * radius, -radius),
                   density=Texture3D(grid=fbm_grid(8, 11) * sigma_s + 0.05),
                   albedo=Texture3D(value=(1.0, 1.0, 1.0)), g3=Texture3D(value=(0.5, 0.5, 0.5)))
    else:
        m = Medium(kind=capi.MEDIUM_HOMOGENEOUS, sigma_a=(0.0, 0.0, 0.0), sigma_s=(sigma_s, sigma_s, sigma_s), g=0.4)
    mi = s.add_medium(m)
    # the normal side of the inward-facing shell is its inside
    s.add(Entity(capi.GEOM_TRIANGLE_BVH, black, emit=(1.0, 1.0, 1.0), tris=tris, med_out=mi))
    cam = {"pos": (0.0, 0.0, 0.0), "dst": (0.0, 1.0, 0.0), "up": (0, 0, 1), "fov": 70.0}
    return s, cam, {"min_depth": 5, "max_depth": 40, "cont_prob": 0.9}


def absorbing_slab(sigma_a=(0.3, 0.6, 1.2), thickness: float = 1.5) -> tuple[Scene, dict, dict]:
    """An emitting quad (radiance 2) seen through a slab between two invisible_surface
    quads filled with a purely absorbing medium: the image is 2 * exp(-sigma_a * L),
    L = thickness / cos(angle of the camera ray to the slab normal)."""
    s = Scene(aggregate=capi.AGGREGATE_NATIVE)
    inv = s.add_material(Material(kind=capi.MAT_INVISIBLE_SURFACE))
    black = s.add_material(Material(kind=capi.MAT_IDEAL_BLACK))
    mi = s.add_medium(Medium(kind=capi.MEDIUM_HOMOGENEOUS, sigma_a=tuple(sigma_a), sigma_s=(0.0, 0.0, 0.0), g=0.0))
    big = 50.0
    y0, y1 = 2.0, 2.0 + thickness
    # front face: normal -y (towards the camera), the medium on its back side
    # (a quad's geometric normal is (B - A) x (C - A))
    s.add(Entity(capi.GEOM_QUAD, inv, points=((-big, y0, -big), (big, y0, -big), (big, y0, big), (-big, y0, big)),
                 med_in=mi))
    # back face: normal +y, the medium on its back side
    s.add(Entity(capi.GEOM_QUAD, inv, points=((-big, y1, -big), (-big, y1, big), (big, y1, big), (big, y1, -big)),
                 med_in=mi))
    # emitter facing the camera (normal -y)
    s.add(Entity(capi.GEOM_QUAD, black, emit=(2.0, 2.0, 2.0),
                 points=((-big, 8.0, -big), (big, 8.0, -big), (big, 8.0, big), (-big, 8.0, big))))
    cam = {"pos": (0.0, 0.0, 0.0), "dst": (0.0, 1.0, 0.0), "up": (0, 0, 1), "fov": 30.0}
    return s, cam, {"min_depth": 8, "max_depth": 8, "cont_prob": 1.0}


def _room_quads(s: Scene, lo, hi, m: int, skip=()) -> None:
    """Inward-facing quads of the box [lo, hi] (names: floor, ceiling, x0, x1, y0, y1)."""
    (x0, y0, z0), (x1, y1, z1) = lo, hi
    faces = {
        "floor": ((x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0)),     # +z
        "ceiling": ((x0, y0, z1), (x0, y1, z1), (x1, y1, z1), (x1, y0, z1)),   # -z
        "x0": ((x0, y0, z0), (x0, y1, z0), (x0, y1, z1), (x0, y0, z1)),        # +x
        "x1": ((x1, y0, z0), (x1, y0, z1), (x1, y1, z1), (x1, y1, z0)),        # -x
        "y0": ((x0, y0, z0), (x0, y0, z1), (x1, y0, z1), (x1, y0, z0)),        # +y
        "y1": ((x0, y1, z0), (x1, y1, z0), (x1, y1, z1), (x0, y1, z1)),        # -y
    }
    for k, q in faces.items():
        if k not in skip:
            s.add(Entity(capi.GEOM_QUAD, m, points=q))


def guided_room(n_clutter_tris: int = 500_000, seed: int = 7, gap: float = 0.05) -> tuple[Scene, dict, dict]:
    """Hard indirect lighting: a closed 10 x 10 x 4 room of Lambert walls (albedo
    0.8) whose only light is a 0.5 x 0.5 emitter in a second room behind a door
    gap `gap` wide; seeded clutter of displaced icospheres and boxes
    (n_clutter_tris triangles) on the floor of the first room.  Light reaches
    the first room only through the gap."""
    rng = np.random.default_rng(seed)
    s = Scene(aggregate=capi.AGGREGATE_BVH)
    wall = s.add_material(Material(color=(0.8, 0.8, 0.8)))
    # room A [0,10]x[0,10]x[0,4] without its x = 10 wall; room B [10,13]x[0,10]x[0,4] without its x = 10 wall
    _room_quads(s, (0, 0, 0), (10, 10, 4), wall, skip=("x1",))
    _room_quads(s, (10, 0, 0), (13, 10, 4), wall, skip=("x0",))
    # the dividing wall at x = 10, two-sided (one quad per side), with a door gap
    # y in [5 - gap/2, 5 + gap/2], z in [0, 2.5]
    g0, g1, dh = 5 - gap / 2, 5 + gap / 2, 2.5
    def wall_pieces(x, sign):
        pieces = [((0, 0), (g0, 4)), ((g1, 0), (10, 4)), ((g0, dh), (g1, 4))]
        for (ya, za), (yb, zb) in pieces:
            if sign > 0:   # normal -x (faces room A)
                q = ((x, ya, za), (x, ya, zb), (x, yb, zb), (x, yb, za))
            else:          # normal +x (faces room B)
                q = ((x, ya, za), (x, yb, za), (x, yb, zb), (x, ya, zb))
            s.add(Entity(capi.GEOM_QUAD, wall, points=q))
    wall_pieces(10.0, +1)
    wall_pieces(10.0, -1)
    # 0.5 x 0.5 light on room B's ceiling near its far wall, facing down
    black = s.add_material(Material(kind=capi.MAT_IDEAL_BLACK))
    lz = 4 - 1e-3
    s.add(Entity(capi.GEOM_QUAD, black, emit=(40.0, 38.0, 34.0),
                 points=((12.0, 2.0, lz), (12.0, 2.5, lz), (12.5, 2.5, lz), (12.5, 2.0, lz))))
    # clutter: displaced icospheres (subdivision 4, 5120 triangles) and 12-triangle boxes
    clutter_mat = [s.add_material(Material(color=tuple(rng.uniform(0.2, 0.8, 3)))) for _ in range(8)]
    pos, faces = icosphere(4)
    nrm = pos / np.linalg.norm(pos, axis=1, keepdims=True)
    tris_left = n_clutter_tris
    k = 0
    while tris_left >= 12:
        c = np.array([rng.uniform(0.5, 9.0), rng.uniform(0.5, 9.5), 0.0])
        m = clutter_mat[k % len(clutter_mat)]
        k += 1
        if tris_left >= len(faces) and rng.random() < 0.8:
            r = rng.uniform(0.15, 0.45)
            disp = 1 + 0.1 * _value_noise(pos * 3 + rng.uniform(0, 100, 3), seed + k)
            v = nrm * (r * disp)[:, None] + c + np.array([0, 0, r])
            s.add(Entity(capi.GEOM_TRIANGLE_BVH, m, tris=mesh_from_indexed(v, vertex_normals(v, faces), faces)))
            tris_left -= len(faces)
        else:
            h = rng.uniform(0.1, 0.8)
            w = rng.uniform(0.1, 0.6, 2)
            s.add(Entity(capi.GEOM_TRIANGLE_BVH, m, tris=box_mesh(c - np.array([w[0], w[1], 0]) / 2,
                                                                  c + np.array([w[0], w[1], 2 * h]) / 2,
                                                                  rng.uniform(0, 90))))
            tris_left -= 12
    cam = {"pos": (1.0, 2.0, 2.6), "dst": (9.0, 6.5, 1.0), "up": (0, 0, 1), "fov": 60.0}
    return s, cam, {"min_depth": 5, "max_depth": 10, "cont_prob": 0.9}


def disney_spheres() -> tuple[Scene, dict, dict]:
    """A 3x3 grid of Disney spheres (metallic, rough, anisotropic, sheen, clearcoat,
    rough transmission) on a Lambert ground, lit by a quad light and a native_sky;
    two of them are triangle meshes so that shading normals differ from geometric ones."""
    s = Scene(aggregate=capi.AGGREGATE_BVH)
    ground = s.add_material(Material(color=(0.6, 0.6, 0.6)))
    s.add(Entity(capi.GEOM_QUAD, ground, points=((-20, -20, 0), (20, -20, 0), (20, 20, 0), (-20, 20, 0))))
    D = capi.MAT_DISNEY
    mats = [
        Material(kind=D, color=(0.8, 0.2, 0.2), roughness=0.5),
        Material(kind=D, color=(0.9, 0.7, 0.3), metallic=1.0, roughness=0.25),
        Material(kind=D, color=(0.3, 0.5, 0.9), metallic=0.5, roughness=0.4, anisotropic=0.7),
        Material(kind=D, color=(0.5, 0.8, 0.4), roughness=0.8, sheen=1.0, sheen_tint=0.5),
        Material(kind=D, color=(0.1, 0.1, 0.1), roughness=0.3, clearcoat=1.0, clearcoat_gloss=0.8),
        Material(kind=D, color=(0.95, 0.95, 0.95), roughness=0.15, transmission=0.9, ior=1.5,
                 transmission_roughness=0.2),
        Material(kind=D, color=(0.7, 0.7, 0.9), roughness=0.35, specular_tint=0.8, color2=(0.6, 0.6, 0.6)),
        Material(kind=D, color=(0.9, 0.5, 0.6), roughness=0.6, transmission=0.5, ior=1.33, metallic=0.2),
        Material(kind=D, color=(0.4, 0.4, 0.4), roughness=0.05, metallic=0.8, clearcoat=0.5, anisotropic=0.3),
    ]
    pos, faces = icosphere(3)
    for k, m in enumerate(mats):
        mi = s.add_material(m)
        cx, cy = (k % 3 - 1) * 2.5, (k // 3 - 1) * 2.5
        if k in (2, 5):
            v = pos * 1.0 + np.array([cx, cy, 1.0])
            s.add(Entity(capi.GEOM_TRIANGLE_BVH, mi, tris=mesh_from_indexed(v, pos, faces)))
        else:
            s.add(Entity(capi.GEOM_SPHERE, mi, radius=1.0, local_to_world=(1, 0, 0, cx, 0, 1, 0, cy, 0, 0, 1, 1.0)))
    light = s.add_material(Material(kind=capi.MAT_IDEAL_BLACK))
    s.add(Entity(capi.GEOM_QUAD, light, emit=(12.0, 11.0, 9.0),
                 points=((-2, -2, 8), (-2, 2, 8), (2, 2, 8), (2, -2, 8))))
    s.env = {"top": (0.4, 0.5, 0.7), "bottom": (0.1, 0.1, 0.1)}
    cam = {"pos": (0.0, -11.0, 7.0), "dst": (0.0, 0.0, 0.8), "up": (0, 0, 1), "fov": 45.0}
    return s, cam, {"min_depth": 3, "max_depth": 8, "cont_prob": 0.9}


SUN_ANGULAR_RADIUS = math.radians(0.265)   # the sun's 0.53 degree disk


def procedural_spheres(n_spheres: int = 200, subdiv: int = 4, seed: int = 1, extent: float = 50.0,
                       ground: bool = True, sun: bool = False) -> tuple[Scene, dict, dict]:
    """Seeded procedural mesh scene: n icospheres (20*4^subdiv triangles each) of
    radius U[0.2, 2] in a [0, extent]^3 box, radially displaced by value noise of
    amplitude 0.1 r, on a ground quad; half Lambert (albedo U[0.1, 0.9]), half
    GGX conductor (roughness U[0.05, 0.6], gold or copper); native_sky.
    200 spheres at subdivision 4 give 1,024,000 triangles.

    sun=True adds the sun as Atrc can express it: a distant emissive disk (Atrc
    has no directional light) 40 scene extents away at elevation 55 degrees,
    facing the scene, subtending 0.53 degrees, with radiance chosen so that its
    irradiance on the ground is about three times the sky's."""
    rng = np.random.default_rng(seed)
    base_v, faces = icosphere(subdiv)
    s = Scene(aggregate=capi.AGGREGATE_BVH)
    # physical (eta, k) of gold and copper at R, G, B; Atrc's metal reads the JSON
    # "eta" as k and "k" as eta (material_creators.cpp:252-266), so they are stored swapped
    gold = ((0.143, 0.374, 1.442), (3.983, 2.386, 1.603))
    copper = ((0.200, 0.924, 1.102), (3.912, 2.452, 2.142))
    for i in range(n_spheres):
        r = rng.uniform(0.2, 2.0)
        c = rng.uniform(0.0, extent, 3)
        disp = 1.0 + 0.1 * _value_noise(base_v * 3.0 + rng.uniform(0, 100, 3), seed * 1000 + i)
        pos = base_v * (r * disp)[:, None] + c
        nor = vertex_normals(pos, faces)
        if rng.uniform() < 0.5:
            m = s.add_material(Material(kind=capi.MAT_IDEAL_DIFFUSE, color=tuple(rng.uniform(0.1, 0.9, 3))))
        else:
            phys_eta, phys_k = gold if rng.uniform() < 0.5 else copper
            m = s.add_material(Material(kind=capi.MAT_METAL, color=(1.0, 1.0, 1.0), eta=phys_k, k=phys_eta,
                                        roughness=float(rng.uniform(0.05, 0.6)), anisotropic=0.0))
        s.add(Entity(capi.GEOM_TRIANGLE_BVH, m, tris=mesh_from_indexed(pos, nor, faces)))
    if ground:
        g = s.add_material(Material(color=(0.5, 0.5, 0.5)))
        lo, hi, z = -100.0, extent + 100.0, -2.5
        s.add(Entity(capi.GEOM_QUAD, g, points=((lo, lo, z), (hi, lo, z), (hi, hi, z), (lo, hi, z))))
    s.env = {"top": (0.9, 0.95, 1.2), "bottom": (0.25, 0.22, 0.2)}
    if sun:
        el, az = math.radians(55.0), math.radians(30.0)
        to_sun = np.array([math.cos(el) * math.cos(az), math.cos(el) * math.sin(az), math.sin(el)])
        dist = 40.0 * extent
        centre = np.full(3, extent / 2) + dist * to_sun
        # disk frame: local +z (its emitting side) turned towards the scene
        z = -to_sun
        x = np.cross(np.array([0.0, 0.0, 1.0]), z)
        x /= np.linalg.norm(x)
        y = np.cross(z, x)
        l2w = np.concatenate([np.stack([x, y, z], axis=1), centre[:, None]], axis=1)
        omega = math.pi * math.tan(SUN_ANGULAR_RADIUS) ** 2
        radiance = 3.0 * math.pi / (omega * math.sin(el))
        black = s.add_material(Material(kind=capi.MAT_IDEAL_BLACK))
        s.add(Entity(capi.GEOM_DISK, black, emit=(radiance, 0.95 * radiance, 0.85 * radiance),
                     radius=dist * math.tan(SUN_ANGULAR_RADIUS), local_to_world=tuple(float(v) for v in l2w.ravel())))
    camera = {"pos": (-25.0, -35.0, 45.0), "dst": (25.0, 25.0, 15.0), "up": (0, 0, 1), "fov": 40.0}
    trace = {"min_depth": 5, "max_depth": 8, "cont_prob": 0.9}
    return s, camera, trace


def batch_scene(seed: int) -> tuple[Scene, dict, dict, int]:
    """One scene of the batched workload: procedural_spheres with a seeded
    triangle count log-uniform in [1e4, 2e6] (subdivision-4 spheres, or
    subdivision 3 below 51,200 triangles) in a box that grows with the count,
    a resolution drawn from {256, 512, 1024} squared, and 32 spp.  Returns
    (scene, camera, trace, resolution)."""
    rng = np.random.default_rng(seed)
    tris = math.exp(rng.uniform(math.log(1e4), math.log(2e6)))
    subdiv = 4 if tris >= 51_200 else 3
    n = max(2, int(round(tris / (20 * 4 ** subdiv))))
    extent = 50.0 * (n / 200.0) ** (1.0 / 3.0)
    s, cam, tr = procedural_spheres(n_spheres=n, subdiv=subdiv, seed=seed, extent=extent)
    k = extent / 50.0
    cam = {"pos": (-25.0 * k, -35.0 * k, 45.0 * k), "dst": (25.0 * k, 25.0 * k, 15.0 * k), "up": (0, 0, 1),
           "fov": 40.0}
    res = int(rng.choice([256, 512, 1024]))
    return s, cam, tr, res


def volumetric_room(n_clutter_tris: int = 500_000, grid_n: int = 256, seed: int = 11, density_max: float = 2.0,
                    slab: bool = True) -> tuple[Scene, dict, dict]:
    """The door-gap room (guided_room) with its first room filled with
    participating media: a heterogeneous medium on a seeded grid_n^3 fBm density
    (4 octaves, max density_max, FP16 values, albedo 0.9, HG g = 0.3) over the whole room, and
    (with slab) a homogeneous slab (sigma_t 0.5, albedo 0.9) between z = 2.0 and 2.5 bounded
    by an invisible_surface box.  Every surface facing the first room has the
    fog on that side (entity med_out, or med_in for the slab's box)."""
    s, cam, tr = guided_room(n_clutter_tris=n_clutter_tris, seed=7)
    # densities rounded to FP16 values: the library stores such a grid in FP16
    density = (fbm_grid(grid_n, seed) * density_max).astype(np.float16).astype(np.float32)
    fog = s.add_medium(Medium(kind=capi.MEDIUM_HETEROGENEOUS,
                              local_to_world=(10.0, 0, 0, 0, 0, 10.0, 0, 0, 0, 0, 4.0, 0),
                              density=Texture3D(grid=density),
                              albedo=Texture3D(value=(0.9, 0.9, 0.9)),
                              g3=Texture3D(value=(0.3, 0.3, 0.3))))
    slab_medium = s.add_medium(Medium(kind=capi.MEDIUM_HOMOGENEOUS, sigma_a=(0.05, 0.05, 0.05),
                                      sigma_s=(0.45, 0.45, 0.45), g=0.0)) if slab else -1
    for e in s.entities:
        if e.tris is not None:                       # clutter: outward normals, fog outside
            e.med_out = fog
            continue
        if any(c != 0 for c in e.emit):              # the light, in the second room
            continue
        p = np.asarray(e.points, np.float64)
        n = np.cross(p[1] - p[0], p[2] - p[0])
        centre = p.mean(axis=0)
        # inward-facing walls of the first room, and the dividing wall's side facing it
        if centre[0] < 10.0 - 1e-6 or (abs(centre[0] - 10.0) < 1e-6 and n[0] < 0):
            e.med_out = fog
    if slab:
        inv = s.add_material(Material(kind=capi.MAT_INVISIBLE_SURFACE))
        s.add(Entity(capi.GEOM_TRIANGLE_BVH, inv, tris=box_mesh((0.5, 0.5, 2.0), (9.5, 9.5, 2.5)),
                     med_in=slab_medium, med_out=fog))
    return s, cam, dict(tr, max_depth=16)


def half_plane_emitter() -> tuple[Scene, dict, dict]:
    """A quad of radiance 1 covering the half-plane x >= 0 of the plane y = 0,
    facing a pinhole camera at (0, -5, 0) that looks along +y: the edge runs
    through the centre of the film, so with an even width it lies exactly on
    the pixel boundary W/2 and a film filter turns it into its own integral."""
    s = Scene(aggregate=capi.AGGREGATE_NATIVE)
    black = s.add_material(Material(kind=capi.MAT_IDEAL_BLACK))
    s.add(Entity(capi.GEOM_QUAD, black, emit=(1.0, 1.0, 1.0),
                 points=((0.0, 0.0, -100.0), (100.0, 0.0, -100.0), (100.0, 0.0, 100.0), (0.0, 0.0, 100.0))))
    cam = {"pos": (0.0, -5.0, 0.0), "dst": (0.0, 0.0, 0.0), "up": (0, 0, 1), "fov": 40.0}
    return s, cam, {"min_depth": 5, "max_depth": 1, "cont_prob": 0.9}


def _inward_shell(radius: float, subdiv: int) -> np.ndarray:
    """Icosphere of the given radius with inward winding and flat inward normals
    (geometric = shading normal), as a (3n, 8) mesh."""
    pos, faces = icosphere(subdiv)
    pos = pos * radius
    faces = faces[:, ::-1].copy()
    a, b, c = pos[faces[:, 0]], pos[faces[:, 1]], pos[faces[:, 2]]
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    tris = mesh_from_indexed(pos, np.zeros_like(pos), faces)
    tris[:, 3:6] = np.repeat(n, 3, axis=0)
    return tris


def emitting_enclosure(kind: str = "quads", rho: float = 0.5, emit: float = 1.0,
                       max_depth: int = 24, outward_double_sided: bool = False) -> tuple[Scene, dict, dict, float]:
    """Closed emitting Lambert enclosure (surface furnace): every wall is a diffuse
    area light of radiance `emit` with a Lambert albedo `rho`.  The camera inside
    sees L = emit * (1 - rho^(D+1)) / (1 - rho) at every pixel for max depth D
    (emission at depth 1, then rho^k from the direct light at the k-th vertex).
    kind = "quads": six inward quads of a 4 x 3 x 2.5 box (quad light sampling);
    "mesh": an inward icosphere of radius 2 (triangle-mesh light sampling).
    Returns (scene, camera, trace settings, expected radiance)."""
    s = Scene(aggregate=capi.AGGREGATE_BVH)
    m = s.add_material(Material(color=(rho, rho, rho)))
    e = (emit, emit, emit)
    if kind == "quads":
        n0 = len(s.entities)
        _room_quads(s, (0.0, 0.0, 0.0), (4.0, 3.0, 2.5), m)
        for ent in s.entities[n0:]:
            ent.emit = e
        cam = {"pos": (1.0, 0.8, 1.0), "dst": (3.0, 2.2, 1.5), "up": (0, 0, 1), "fov": 70.0}
    else:
        s.add(Entity(capi.GEOM_TRIANGLE_BVH, m, emit=e, tris=_inward_shell(2.0, 2)))
        cam = {"pos": (0.2, -0.3, 0.1), "dst": (1.0, 1.0, 0.5), "up": (0, 0, 1), "fov": 70.0}
    if outward_double_sided:
        # the same walls facing outwards, made double-sided: the inside sees
        # their back sides, which shade and emit like front sides
        for ent in s.entities:
            ent.double_sided = True
            if ent.tris is not None:
                t = ent.tris.reshape(-1, 3, 8)[:, ::-1].reshape(-1, 8).copy()
                t[:, 3:6] *= -1
                ent.tris = t
            else:
                ent.points = tuple(reversed(ent.points))
    expected = emit * (1.0 - rho ** (max_depth + 1)) / (1.0 - rho)
    return s, cam, {"min_depth": 3, "max_depth": max_depth, "cont_prob": 0.8}, expected


def light_in_shell(kind: str = "sphere", rho: float = 0.6, emit: float = 3.0, shell_radius: float = 4.0,
                   subdiv: int = 6, double_sided: bool = False, face_up: bool = False) -> tuple[Scene, dict, dict, float]:
    """Direct light (max depth 1) on an inward Lambert shell of radius D from an
    emitter whose irradiance is the same at every shell point, so that every pixel
    that sees the shell has one closed-form value:
    * "sphere": a sphere light of radius R = 1 at the centre; E = pi L (R/D)^2,
      radiance rho L (R/D)^2.
    * "disk": a disk light facing -z (or +z with face_up, then seen from below
      only through double_sided) whose rim lies on the sphere at height
      z0 = D/2.  By the sphere's configuration-factor theorem (from a point on a
      sphere, any area on the sphere has form factor A / (4 pi D^2)) and because
      the disk and the spherical cap above it share their boundary, every shell
      point below the disk sees it with form factor (D - z0) / (2 D): radiance
      rho L (D - z0) / (2 D).
    The shell's flat faces lie within 5e-5 D of the sphere at subdivision 6.
    Returns (scene, camera, trace settings, expected radiance)."""
    s = Scene(aggregate=capi.AGGREGATE_BVH)
    m = s.add_material(Material(color=(rho, rho, rho)))
    black = s.add_material(Material(kind=capi.MAT_IDEAL_BLACK))
    D = shell_radius
    s.add(Entity(capi.GEOM_TRIANGLE_BVH, m, tris=_inward_shell(D, subdiv)))
    e = (emit, emit, emit)
    if kind == "sphere":
        R = 1.0
        s.add(Entity(capi.GEOM_SPHERE, black, emit=e, radius=R, double_sided=double_sided))
        expected = rho * emit * (R / D) ** 2
        cam = {"pos": (0.0, -1.6, -0.4), "dst": (0.3, -4.0, -1.0), "up": (0, 0, 1), "fov": 60.0}
    else:
        z0 = D / 2
        a = math.sqrt(D * D - z0 * z0)
        # rotation by pi about x: the disk's local +z normal points down
        l2w = (1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, z0) if face_up else (1, 0, 0, 0, 0, -1, 0, 0, 0, 0, -1, z0)
        s.add(Entity(capi.GEOM_DISK, black, emit=e, radius=a, local_to_world=l2w, double_sided=double_sided))
        expected = rho * emit * (D - z0) / (2 * D)
        cam = {"pos": (0.0, 0.0, 0.5), "dst": (0.4, 0.3, -3.0), "up": (0, 1, 0), "fov": 70.0}
    return s, cam, {"min_depth": 5, "max_depth": 1, "cont_prob": 0.9}, expected


def random_rays(scene_bounds: tuple, n: int, seed: int, inside: bool = True):
    """Seeded rays: origins uniform in the box, directions uniform on the sphere."""
    rng = np.random.default_rng(seed)
    lo, hi = (np.asarray(x, np.float64) for x in scene_bounds)
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o.astype(np.float32), d.astype(np.float32)


# ---------------------------------------------------------------------------------- Atrc JSON

def _texture(c):
    c = [float(x) for x in (c if isinstance(c, (tuple, list, np.ndarray)) else (c,))]
    return {"type": "constant", "texel": c}


def _material_json(m: Material) -> dict:
    if m.kind == capi.MAT_IDEAL_DIFFUSE:
        return {"type": "ideal_diffuse", "albedo": _texture(m.color)}
    if m.kind == capi.MAT_METAL:
        return {"type": "metal", "color": _texture(m.color), "eta": _texture(m.eta), "k": _texture(m.k),
                "roughness": _texture(m.roughness), "anisotropic": _texture(m.anisotropic)}
    if m.kind == capi.MAT_GLASS:
        out = {"type": "glass", "ior": _texture(m.ior)}
        if m.separate_refraction:
            out["color_reflection_map"] = _texture(m.color)
            out["color_refraction_map"] = _texture(m.color2)
        else:
            out["color_map"] = _texture(m.color)
        return out
    if m.kind == capi.MAT_MIRROR:
        return {"type": "mirror", "color_map": _texture(m.color), "eta": _texture(m.eta), "k": _texture(m.k)}
    return {"type": "ideal_black"}


def _transform_json(l2w) -> list:
    m = np.array(l2w, np.float64).reshape(3, 4)
    return [{"type": "translate", "offset": [float(x) for x in m[:, 3]]},
            {"type": "mat3", "row_major": [float(x) for x in m[:, :3].reshape(-1)]}]


def to_atrc_json(scene: Scene, camera: dict, trace: dict, width: int, height: int, spp: int, mesh_dir,
                 renderer: str = "gpu_pt", seed: int = 42, post_processors: list | None = None,
                 eps: float | None = None, renderer_extra: dict | None = None) -> str:
    """Atrc JSON ({"scene", "rendering"}) for a Python-built scene; meshes are
    written as .bm files into mesh_dir and referenced via ${scene-directory}."""
    import json
    from pathlib import Path

    mesh_dir = Path(mesh_dir)
    ents = []
    for i, e in enumerate(scene.entities):
        if e.geometry == capi.GEOM_QUAD:
            g = {"type": "quad", **{k: [float(x) for x in p] for k, p in zip("ABCD", e.points)}}
            for k, t in zip(("tA", "tB", "tC", "tD"), e.uvs):
                g[k] = [float(x) for x in t]
        elif e.geometry == capi.GEOM_TRIANGLE_BVH:
            name = f"mesh_{i}.bm"
            write_bin_mesh(mesh_dir / name, e.tris)
            g = {"type": "triangle_bvh", "filename": "${scene-directory}/" + name}
        elif e.geometry == capi.GEOM_TRIANGLE:
            g = {"type": "triangle", **{k: [float(x) for x in p] for k, p in zip("ABC", e.points)},
                 "transform": _transform_json(e.local_to_world)}
        else:
            g = {"type": "sphere" if e.geometry == capi.GEOM_SPHERE else "disk", "radius": float(e.radius),
                 "transform": _transform_json(e.local_to_world)}
        if e.double_sided:
            g = {"type": "double_sided", "internal": g}
        ent = {"type": "geometric", "geometry": g, "material": _material_json(scene.materials[e.material])}
        if any(e.emit):
            ent["emit_radiance"] = [float(x) for x in e.emit]
        if e.no_denoise:
            ent["no_denoise"] = 1
        ents.append(ent)
    sc = {"type": "default", "entities": ents}
    if scene.env is not None:
        sc["env"] = {"type": "native_sky", "top": [float(x) for x in scene.env["top"]],
                     "bottom": [float(x) for x in scene.env["bottom"]]}
    sc["aggregate"] = ({"type": "bvh", "max_leaf_size": scene.max_leaf_size}
                       if scene.aggregate == capi.AGGREGATE_BVH else {"type": "native"})
    cam = {"type": "thin_lens", "pos": [float(x) for x in camera["pos"]], "dst": [float(x) for x in camera["dst"]],
           "up": [float(x) for x in camera.get("up", (0, 0, 1))], "fov": float(camera.get("fov", 40.0))}
    rend = {"type": renderer, "spp": spp, "min_depth": trace.get("min_depth", 5),
            "max_depth": trace.get("max_depth", 10), "cont_prob": trace.get("cont_prob", 0.9), "seed": seed,
            **(renderer_extra or {})}
    rendering = {"width": width, "height": height, "camera": cam, "renderer": rend, "reporter": {"type": "noout"}}
    if post_processors:
        rendering["post_processors"] = post_processors
    if eps is not None:
        rendering["eps"] = eps
    return json.dumps({"scene": sc, "rendering": rendering}, indent=1)


def write_bin_mesh(path, tris: np.ndarray) -> None:
    """Atrc's .bm format: uint64 triangle count, then 3 vertices of 8 floats each."""
    tris = np.ascontiguousarray(tris, np.float32)
    with open(path, "wb") as f:
        f.write(np.uint64(len(tris) // 3).tobytes())
        f.write(tris.tobytes())
