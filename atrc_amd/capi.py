"""ctypes mirror of ``include/atrc_gpu.h`` and a thin Python wrapper.

The structures below must match the C header field for field; the CPU test
suite checks their sizes against the compiled library.  The wrapper raises
``AgxError`` with ``agx_last_error()`` on any non-zero status, and loading
fails loudly when ``libatrc_gpu.so`` has not been built: there is no fallback
path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
GPU_LIB_PATH = PKG / "libatrc_gpu.so"

AGX_ABI_VERSION = 13

AGX_OK, AGX_E_INVALID_ARG, AGX_E_HIP, AGX_E_NO_SCENE, AGX_E_UNSUPPORTED, AGX_E_CANCELLED, \
    AGX_E_OUT_OF_MEMORY, AGX_E_LIMIT = range(8)

GEOM_TRIANGLE_BVH, GEOM_QUAD, GEOM_TRIANGLE, GEOM_SPHERE, GEOM_DISK = range(5)
MEDIUM_VOID, MEDIUM_HOMOGENEOUS, MEDIUM_HETEROGENEOUS = range(3)
TEX3D_CONSTANT, TEX3D_GRID = range(2)
WRAP_CLAMP, WRAP_REPEAT, WRAP_MIRROR = range(3)
MAT_IDEAL_DIFFUSE, MAT_METAL, MAT_GLASS, MAT_MIRROR, MAT_IDEAL_BLACK, MAT_INVISIBLE_SURFACE, MAT_DISNEY = range(7)
ENV_NONE, ENV_NATIVE_SKY, ENV_IBL = 0, 1, 2
TEX2D_CONSTANT, TEX2D_IMAGE, TEX2D_CHECKER = 0, 1, 2
FILTER_BOX, FILTER_GAUSSIAN = 0, 1
ACCEL_ATRC, ACCEL_BVH4, ACCEL_BVH4_WATERTIGHT = 0, 1, 2
AGGREGATE_NATIVE, AGGREGATE_BVH = 0, 1

F3 = C.c_float * 3


class agx_vertex(C.Structure):
    _fields_ = [("pos", F3), ("nor", F3), ("uv", C.c_float * 2)]


class agx_entity_desc(C.Structure):
    _fields_ = [
        ("geometry", C.c_int32), ("material", C.c_int32), ("medium_in", C.c_int32), ("medium_out", C.c_int32),
        ("emit_radiance", F3), ("power", C.c_float), ("no_denoise", C.c_int32), ("tri_count", C.c_uint32),
        ("tris", C.POINTER(agx_vertex)), ("p", (C.c_float * 3) * 4), ("t", (C.c_float * 2) * 4),
        ("radius", C.c_float), ("local_to_world", C.c_float * 12), ("world_to_local", C.c_float * 12),
        ("double_sided", C.c_int32),
    ]


class agx_material_desc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("color", F3), ("color2", F3), ("eta", F3), ("k", F3),
        ("roughness", C.c_float), ("anisotropic", C.c_float), ("ior", C.c_float),
        ("glass_separate_refraction", C.c_int32), ("metallic", C.c_float), ("transmission", C.c_float),
        ("transmission_roughness", C.c_float), ("specular_tint", C.c_float), ("sheen", C.c_float),
        ("sheen_tint", C.c_float), ("clearcoat", C.c_float), ("clearcoat_gloss", C.c_float),
    ]


class agx_texture3d_desc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("linear", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
        ("depth", C.c_int32), ("channels", C.c_int32), ("data", C.POINTER(C.c_float)), ("value", F3),
        ("transform", C.c_float * 12), ("wrap", C.c_int32 * 3), ("inv_gamma", C.c_float),
    ]


class agx_medium_desc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("max_scattering_count", C.c_int32), ("sigma_a", F3), ("sigma_s", F3), ("g", C.c_float),
        ("local_to_world", C.c_float * 12), ("world_to_local", C.c_float * 12), ("density", agx_texture3d_desc),
        ("albedo", agx_texture3d_desc), ("g3", agx_texture3d_desc), ("white_for_indirect", C.c_int32),
    ]


class agx_texture2d_desc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("linear", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
        ("data", C.POINTER(C.c_float)), ("value", F3), ("value2", F3), ("grid_count", C.c_int32),
        ("transform", C.c_float * 6), ("wrap", C.c_int32 * 2), ("inv_gamma", C.c_float),
    ]


class agx_env_desc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("top", F3), ("bottom", F3), ("power", C.c_float), ("tex", agx_texture2d_desc),
                ("no_importance_sampling", C.c_int32)]


class agx_camera_desc(C.Structure):
    _fields_ = [
        ("pos", F3), ("dst", F3), ("up", F3), ("fov_rad", C.c_float), ("lens_radius", C.c_float),
        ("focal_distance", C.c_float), ("film_aspect", C.c_float),
    ]


class agx_scene_desc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32), ("entity_count", C.c_uint32), ("entities", C.POINTER(agx_entity_desc)),
        ("material_count", C.c_uint32), ("materials", C.POINTER(agx_material_desc)), ("env", agx_env_desc),
        ("aggregate", C.c_int32), ("max_leaf_size", C.c_int32), ("medium_count", C.c_uint32),
        ("media", C.POINTER(agx_medium_desc)), ("accel", C.c_int32),
    ]


class agx_render_params(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("min_depth", C.c_int32),
        ("max_depth", C.c_int32), ("specular_depth", C.c_int32), ("cont_prob", C.c_float), ("use_mis", C.c_int32),
        ("eps", C.c_float), ("filter_radius", C.c_float), ("seed", C.c_uint64), ("camera", agx_camera_desc),
        ("collect_stats", C.c_int32), ("sample_begin", C.c_int32), ("guide", C.c_int32), ("guide_alpha", C.c_float),
        ("guide_train", C.c_int32), ("guide_batch", C.c_int32), ("guide_steps", C.c_int32), ("guide_lr", C.c_float),
        ("guide_reset", C.c_int32), ("guide_seed", C.c_uint32), ("filter", C.c_int32), ("filter_alpha", C.c_float),
        ("task_grid_size", C.c_int32), ("guide_medium_depth", C.c_int32),
    ]


class agx_stats_t(C.Structure):
    _fields_ = [
        ("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("paths", C.c_uint64),
        ("closest_node_visits", C.c_uint64), ("closest_tri_tests", C.c_uint64), ("closest_hits", C.c_uint64),
        ("shadow_node_visits", C.c_uint64), ("shadow_tri_tests", C.c_uint64),
        ("closest_bvh_bytes", C.c_uint64), ("shadow_bvh_bytes", C.c_uint64),
        ("total_ms", C.c_double), ("trace_closest_ms", C.c_double), ("trace_shadow_ms", C.c_double),
        ("shade_ms", C.c_double), ("trace_closest_launches", C.c_uint64), ("trace_shadow_launches", C.c_uint64),
        ("guide_infer_ms", C.c_double), ("guide_train_ms", C.c_double), ("guide_samples", C.c_uint64),
        ("guide_steps", C.c_uint64), ("guide_rows", C.c_uint64), ("guide_trained", C.c_uint64), ("node_bytes", C.c_uint32), ("prim_bytes", C.c_uint32), ("info_bytes", C.c_uint32), ("reserved", C.c_uint32),
        ("host_ms", C.c_double), ("readback_ms", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class agx_debug_summary(C.Structure):
    _fields_ = [
        ("entities", C.c_uint32), ("lights", C.c_uint32), ("blas_nodes", C.c_uint32), ("blas_prims", C.c_uint32),
        ("tlas_nodes", C.c_uint32), ("max_blas_depth", C.c_uint32), ("max_tlas_depth", C.c_uint32),
        ("device_bytes", C.c_uint64), ("world_centre_x", C.c_float), ("world_centre_y", C.c_float),
        ("world_centre_z", C.c_float), ("world_radius", C.c_float), ("bvh4_nodes", C.c_uint64),
        ("bvh4_leaves", C.c_uint64), ("bvh4_depth", C.c_uint32), ("pad_", C.c_uint32), ("bvh4_sah", C.c_double),
        ("grid_texels_f32", C.c_uint64), ("grid_texels_f16", C.c_uint64),
    ]


# every symbol the public headers declare
class agx_batch_item(C.Structure):
    _fields_ = [
        ("scene", C.POINTER(agx_scene_desc)), ("params", C.POINTER(agx_render_params)), ("rgb", C.POINTER(C.c_float)),
        ("status", C.c_int32), ("worker", C.c_int32), ("rays", C.c_uint64), ("build_ms", C.c_double),
        ("render_ms", C.c_double),
    ]


class agx_batch_report(C.Structure):
    _fields_ = [
        ("workers", C.c_int32), ("tuned", C.c_int32), ("wall_ms", C.c_double), ("rays", C.c_uint64),
        ("tune_rays_per_s", C.c_double * 4), ("pass_paths_log2", C.c_int32), ("trace_block", C.c_int32),
        ("groups", C.c_int32), ("build_threads", C.c_int32), ("build_wall_ms", C.c_double),
        ("render_ms", C.c_double), ("tune_configs", C.c_int32), ("reserved", C.c_int32),
    ]


EXPORTS = [
    "agx_abi_version", "agx_create", "agx_upload_scene", "agx_render", "agx_trace_rays", "agx_get_stats",
    "agx_last_error", "agx_destroy", "agx_debug_triangle_bvh", "agx_debug_scene_summary", "agx_debug_copy_kernel",
    "agx_debug_bvh4", "agx_debug_guide_infer", "agx_debug_guide_train", "agx_debug_replay_rays", "agx_render_batch", "agx_render_batch_graph",
    "agx_debug_guide_grad_mfma", "agx_trace_rays_noalloc",
]
NOALLOC_MAX_RAYS = 131072   # AGX_NOALLOC_MAX_RAYS
GUIDE_PARAMS = 31008   # AGX_GUIDE_PARAMS


class AgxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"agx error {code}: {msg}")
        self.code = code


_lib = None


def load_library() -> C.CDLL:
    """Loads libatrc_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not GPU_LIB_PATH.exists():
        raise RuntimeError(f"{GPU_LIB_PATH} is missing: run `python -m atrc_amd.build gpu` (or __graft_entry__.build())")
    lib = C.CDLL(str(GPU_LIB_PATH))
    P = C.POINTER
    fp = P(C.c_float)
    ip = P(C.c_int32)
    lib.agx_abi_version.restype = C.c_int
    lib.agx_create.argtypes = [C.c_int, P(C.c_void_p)]
    lib.agx_upload_scene.argtypes = [C.c_void_p, P(agx_scene_desc)]
    lib.agx_render.argtypes = [C.c_void_p, P(agx_render_params), P(C.c_int32), fp, fp, fp, fp]
    lib.agx_trace_rays.argtypes = [C.c_void_p, fp, fp, fp, fp, C.c_uint32, C.c_int32, ip, ip, fp, fp, fp]
    lib.agx_trace_rays_noalloc.argtypes = [C.c_void_p, fp, fp, fp, fp, C.c_uint32, C.c_int32, ip, ip, fp, fp, fp]
    lib.agx_get_stats.argtypes = [C.c_void_p, P(agx_stats_t)]
    lib.agx_last_error.restype = C.c_char_p
    lib.agx_destroy.argtypes = [C.c_void_p]
    lib.agx_destroy.restype = None
    lib.agx_debug_triangle_bvh.argtypes = [P(agx_vertex), C.c_uint32, P(C.c_uint32), P(C.c_uint32), fp, fp]
    lib.agx_debug_scene_summary.argtypes = [P(agx_scene_desc), P(agx_debug_summary)]
    lib.agx_debug_copy_kernel.argtypes = [C.c_void_p, C.c_uint64, P(C.c_double)]
    lib.agx_debug_guide_grad_mfma.argtypes = [C.c_void_p, P(C.c_float), P(C.c_float), C.c_uint32, C.c_float,
                                              P(C.c_float)]
    lib.agx_render_batch.argtypes = [C.c_void_p, P(agx_batch_item), C.c_uint32, C.c_int32, P(agx_batch_report)]
    lib.agx_render_batch_graph.argtypes = [C.c_void_p, P(agx_batch_item), C.c_uint32, C.c_int32,
                                           P(agx_batch_report)]
    lib.agx_debug_replay_rays.argtypes = [C.c_void_p, P(agx_render_params), C.c_uint32, C.c_int, P(C.c_uint64),
                                          P(C.c_double)]
    lib.agx_debug_bvh4.argtypes = [P(agx_scene_desc), P(C.c_uint64), P(C.c_uint64), P(C.c_float), P(C.c_float)]
    lib.agx_debug_guide_infer.argtypes = [C.c_void_p, P(C.c_float), P(C.c_float), C.c_uint32, P(C.c_float)]
    lib.agx_debug_guide_train.argtypes = [C.c_void_p, P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_float),
                                          C.c_uint32, C.c_float, C.c_float, C.c_int, P(C.c_float)]
    if lib.agx_abi_version() != AGX_ABI_VERSION:
        raise RuntimeError("libatrc_gpu.so ABI version mismatch")
    _lib = lib
    return lib


def _check(code: int) -> None:
    if code != AGX_OK:
        raise AgxError(code, load_library().agx_last_error().decode())


def fptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def iptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Context:
    """One device context (agx_create / agx_destroy)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        _check(self.lib.agx_create(device, C.byref(h)))
        self.h = h
        self._keep = None

    def close(self) -> None:
        if self.h:
            self.lib.agx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_scene(self, desc: "agx_scene_desc") -> None:
        _check(self.lib.agx_upload_scene(self.h, C.byref(desc)))

    def render(self, params: agx_render_params, cancel: C.c_int32 | None = None, aovs: bool = True,
               out: np.ndarray | None = None) -> dict:
        """agx_render; aovs=False reads back the image only (albedo / normal / denoise stay None).
        out: a C-contiguous float32 (height, width, 3) array the image is written into
        (a render loop reuses it instead of mapping fresh host pages every call)."""
        wh = params.width * params.height
        if out is not None:
            if (out.dtype != np.float32 or out.shape != (params.height, params.width, 3)
                    or not out.flags["C_CONTIGUOUS"]):
                raise ValueError("out must be a C-contiguous float32 array of shape (height, width, 3)")
            rgb = out
        else:
            rgb = np.zeros((params.height, params.width, 3), np.float32)
        alb = np.zeros_like(rgb) if aovs else None
        nor = np.zeros_like(rgb) if aovs else None
        den = np.zeros((params.height, params.width), np.float32) if aovs else None
        code = self.lib.agx_render(self.h, C.byref(params), C.byref(cancel) if cancel is not None else None,
                                   fptr(rgb), fptr(alb), fptr(nor), fptr(den))
        if code not in (AGX_OK, AGX_E_CANCELLED):
            _check(code)
        assert rgb.size == 3 * wh
        return {"image": rgb, "albedo": alb, "normal": nor, "denoise": den, "cancelled": code == AGX_E_CANCELLED}

    def render_batch(self, descs, params, workers: int = 0, graph: bool = False):
        """agx_render_batch over (scene descriptor, render params) pairs: returns
        (images, per-item records, report)."""
        n = len(descs)
        items = (agx_batch_item * n)()
        images = []
        for i, (d, p) in enumerate(zip(descs, params)):
            img = np.zeros((p.height, p.width, 3), np.float32)
            images.append(img)
            items[i].scene = C.pointer(d)
            items[i].params = C.pointer(p)
            items[i].rgb = fptr(img)
        rep = agx_batch_report()
        fn = self.lib.agx_render_batch_graph if graph else self.lib.agx_render_batch
        _check(fn(self.h, items, n, workers, C.byref(rep)))
        recs = [{"status": it.status, "worker": it.worker, "rays": it.rays, "build_ms": it.build_ms,
                 "render_ms": it.render_ms} for it in items]
        report = {"workers": rep.workers, "tuned": bool(rep.tuned), "wall_ms": rep.wall_ms, "rays": rep.rays,
                  "tune_rays_per_s": list(rep.tune_rays_per_s), "pass_paths_log2": rep.pass_paths_log2,
                  "trace_block": rep.trace_block, "groups": rep.groups, "build_threads": rep.build_threads,
                  "build_wall_ms": rep.build_wall_ms, "render_ms": rep.render_ms, "tune_configs": rep.tune_configs}
        return images, recs, report

    def trace_rays(self, o: np.ndarray, d: np.ndarray, tmin: np.ndarray, tmax: np.ndarray, any_hit: bool = False,
                   noalloc: bool = False):
        """Ray queries (agx_trace_rays, or agx_trace_rays_noalloc when noalloc)."""
        n = len(o)
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        tmin = np.ascontiguousarray(tmin, np.float32)
        tmax = np.ascontiguousarray(tmax, np.float32)
        ent = np.zeros(n, np.int32)
        prim = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        fn = self.lib.agx_trace_rays_noalloc if noalloc else self.lib.agx_trace_rays
        _check(fn(self.h, fptr(o), fptr(d), fptr(tmin), fptr(tmax), n, 1 if any_hit else 0,
                  iptr(ent), iptr(prim), fptr(t), fptr(u), fptr(v)))
        return {"entity": ent, "prim": prim, "t": t, "u": u, "v": v}

    def guide_infer(self, params: np.ndarray, inputs: np.ndarray) -> np.ndarray:
        """Path-guiding network on (n, 8) inputs -> (n, 40): 8 x (mu.xyz, kappa), 8 weights."""
        params = np.ascontiguousarray(params, np.float32)
        inputs = np.ascontiguousarray(inputs, np.float32)
        assert params.size == GUIDE_PARAMS and inputs.ndim == 2 and inputs.shape[1] == 8
        out = np.zeros((len(inputs), 40), np.float32)
        _check(self.lib.agx_debug_guide_infer(self.h, fptr(params), fptr(inputs), len(inputs), fptr(out)))
        return out

    def guide_train(self, params, m, v, samples, wscale: float, lr: float, t: int):
        """One training step; returns (params, m, v, gradient)."""
        params = np.array(params, np.float32)
        m = np.array(m, np.float32)
        v = np.array(v, np.float32)
        samples = np.ascontiguousarray(samples, np.float32)
        assert samples.ndim == 2 and samples.shape[1] == 12
        g = np.zeros(GUIDE_PARAMS, np.float32)
        _check(self.lib.agx_debug_guide_train(self.h, fptr(params), fptr(m), fptr(v), fptr(samples), len(samples),
                                              wscale, lr, t, fptr(g)))
        return params, m, v, g

    def guide_grad_mfma(self, params, samples, wscale: float) -> np.ndarray:
        """Gradient of the training loss from the matrix-core training path."""
        params = np.ascontiguousarray(params, np.float32)
        samples = np.ascontiguousarray(samples, np.float32)
        assert samples.ndim == 2 and samples.shape[1] == 12
        g = np.zeros(GUIDE_PARAMS, np.float32)
        _check(self.lib.agx_debug_guide_grad_mfma(self.h, fptr(params), fptr(samples), len(samples), wscale, fptr(g)))
        return g

    def copy_kernel(self, nbytes: int) -> float:
        """One plain copy kernel of nbytes (traffic-counter calibration); device ms."""
        ms = C.c_double()
        _check(self.lib.agx_debug_copy_kernel(self.h, nbytes, C.byref(ms)))
        return ms.value

    def replay_rays(self, params, max_rays: int = 1 << 26, reps: int = 3) -> dict:
        """Renders `params` capturing every ray query, then replays the captured
        closest-hit and any-hit sets with the traversal kernel alone."""
        n = (C.c_uint64 * 2)()
        ms = (C.c_double * 4)()
        _check(self.lib.agx_debug_replay_rays(self.h, C.byref(params), max_rays, reps, n, ms))
        return {"closest_rays": n[0], "shadow_rays": n[1], "replay_closest_ms": ms[0], "replay_shadow_ms": ms[1],
                "render_closest_ms": ms[2], "render_shadow_ms": ms[3]}

    def stats(self) -> dict:
        s = agx_stats_t()
        _check(self.lib.agx_get_stats(self.h, C.byref(s)))
        return s.as_dict()


def debug_triangle_bvh(tris: np.ndarray):
    """Atrc triangle BVH of a mesh (tris: (n*3, 8) float32 vertex records)."""
    lib = load_library()
    tris = np.ascontiguousarray(tris, np.float32)
    n = len(tris) // 3
    cnt = C.c_uint32()
    depth = C.c_uint32()
    vp = tris.ctypes.data_as(C.POINTER(agx_vertex))
    _check(lib.agx_debug_triangle_bvh(vp, n, C.byref(cnt), C.byref(depth), None, None))
    nodes = np.zeros((cnt.value, 8), np.float32)
    prims = np.zeros((n, 12), np.float32)
    _check(lib.agx_debug_triangle_bvh(vp, n, C.byref(cnt), C.byref(depth), fptr(nodes), fptr(prims)))
    return nodes, prims, depth.value


def debug_scene_summary(desc: agx_scene_desc) -> dict:
    lib = load_library()
    s = agx_debug_summary()
    _check(lib.agx_debug_scene_summary(C.byref(desc), C.byref(s)))
    return {name: getattr(s, name) for name, _ in s._fields_}


def debug_bvh4(desc: agx_scene_desc) -> tuple:
    """(nodes (n, 4, 4) float32, prims (m, 3, 4) float32) of an ACCEL_BVH4 or ACCEL_BVH4_WATERTIGHT descriptor."""
    lib = load_library()
    nn, npr = C.c_uint64(0), C.c_uint64(0)
    _check(lib.agx_debug_bvh4(C.byref(desc), C.byref(nn), C.byref(npr), None, None))
    nodes = np.zeros((nn.value, 4, 4), np.float32)
    prims = np.zeros((npr.value, 3, 4), np.float32)
    _check(lib.agx_debug_bvh4(C.byref(desc), C.byref(nn), C.byref(npr), fptr(nodes), fptr(prims)))
    return nodes, prims


# ---------------------------------------------------------------------------------- C++ host
HOST_LIB_PATH = PKG / "libatrc_host.so"


class atrc_host_session(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("eps", C.c_float), ("filter_radius", C.c_float),
                ("camera", agx_camera_desc), ("session_count", C.c_int32), ("filter", C.c_int32),
                ("filter_alpha", C.c_float)]


_host = None


def load_host() -> C.CDLL:
    global _host
    if _host is not None:
        return _host
    if not HOST_LIB_PATH.exists():
        raise RuntimeError(f"{HOST_LIB_PATH} is missing: run `python -m atrc_amd.build host`")
    lib = C.CDLL(str(HOST_LIB_PATH))
    lib.atrc_host_load.restype = C.c_void_p
    lib.atrc_host_load.argtypes = [C.c_char_p, C.c_char_p]
    lib.atrc_host_scene_desc.restype = C.POINTER(agx_scene_desc)
    lib.atrc_host_scene_desc.argtypes = [C.c_void_p]
    lib.atrc_host_session_info.argtypes = [C.c_void_p, C.c_int, C.POINTER(atrc_host_session)]
    fp = C.POINTER(C.c_float)
    lib.atrc_host_execute.argtypes = [C.c_void_p, C.c_int, fp, fp, fp, fp]
    lib.atrc_host_free.argtypes = [C.c_void_p]
    lib.atrc_host_free.restype = None
    lib.atrc_host_last_error.restype = C.c_char_p
    _host = lib
    return lib


class HostScene:
    """A scene and its render sessions created by the C++ host from Atrc JSON."""

    def __init__(self, json_text: str, scene_dir: str):
        self.lib = load_host()
        self.h = self.lib.atrc_host_load(json_text.encode(), str(scene_dir).encode())
        if not self.h:
            raise RuntimeError("atrc host: " + self.lib.atrc_host_last_error().decode())

    def desc(self) -> agx_scene_desc:
        return self.lib.atrc_host_scene_desc(self.h).contents

    def session(self, i: int = 0) -> atrc_host_session:
        s = atrc_host_session()
        if self.lib.atrc_host_session_info(self.h, i, C.byref(s)):
            raise RuntimeError(self.lib.atrc_host_last_error().decode())
        return s

    def execute(self, i: int = 0) -> dict:
        s = self.session(i)
        rgb = np.zeros((s.height, s.width, 3), np.float32)
        alb = np.zeros_like(rgb)
        nor = np.zeros_like(rgb)
        den = np.zeros((s.height, s.width), np.float32)
        if self.lib.atrc_host_execute(self.h, i, fptr(rgb), fptr(alb), fptr(nor), fptr(den)):
            raise RuntimeError("atrc host: " + self.lib.atrc_host_last_error().decode())
        return {"image": rgb, "albedo": alb, "normal": nor, "denoise": den}

    def close(self):
        if self.h:
            self.lib.atrc_host_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
