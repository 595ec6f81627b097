"""ORACLE — test infrastructure only.

ctypes wrapper of liboracle.so (the CPU restatement of Atrc `pt`).  Imported
only by tests, the smoke check and the benchmark's CPU-baseline leg.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_lib = None


def load(path: Path | None = None) -> C.CDLL:
    """The oracle library (liboracle.so), or the library at `path` (for
    instance a -march=native build of the same source, which the module then
    uses from here on)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    lib_path = Path(path) if path is not None else LIB_PATH
    if not lib_path.exists():
        raise RuntimeError(f"{lib_path} missing: run `make -C oracle`")
    lib = C.CDLL(str(lib_path))
    P = C.POINTER
    lib.oracle_last_error.restype = C.c_char_p
    lib.oracle_scene_create.restype = C.c_void_p
    lib.oracle_scene_create.argtypes = [C.c_void_p, C.c_int, C.c_float]
    lib.oracle_scene_destroy.argtypes = [C.c_void_p]
    lib.oracle_trace_rays.argtypes = [C.c_void_p, P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_float), C.c_uint32,
                                      C.c_int, P(C.c_int32), P(C.c_int32), P(C.c_float), P(C.c_float), P(C.c_float),
                                      P(C.c_uint64)]
    lib.oracle_render.argtypes = [C.c_void_p, C.c_void_p, C.c_int, P(C.c_double), P(C.c_double), P(C.c_double),
                                  P(C.c_double), P(C.c_uint64), C.c_int, C.c_int, P(C.c_double)]
    lib.oracle_triangle_soup_closest.argtypes = [P(C.c_float), C.c_uint32, P(C.c_float), P(C.c_float),
                                                 P(C.c_float), P(C.c_float), C.c_uint32, C.c_int, P(C.c_int32),
                                                 P(C.c_float), P(C.c_float), P(C.c_float)]
    lib.oracle_pcg_stream.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, P(C.c_float)]
    lib.oracle_pcg_raw.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, P(C.c_uint32)]
    lib.oracle_distribution.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, P(C.c_double)]
    lib.oracle_bsdf.argtypes = [C.c_void_p, C.c_int, P(C.c_double), P(C.c_double), P(C.c_double), C.c_double,
                                C.c_double, C.c_double, C.c_int, P(C.c_double)]
    lib.oracle_medium.argtypes = [C.c_void_p, C.c_int, C.c_int, P(C.c_double), P(C.c_double), C.c_int, C.c_uint64,
                                  C.c_uint32, P(C.c_double)]
    lib.oracle_tex3d.argtypes = [C.c_void_p, C.c_int, C.c_int, P(C.c_double), C.c_uint32, P(C.c_double)]
    lib.oracle_env.argtypes = [C.c_void_p, C.c_int, P(C.c_double), C.c_uint32, P(C.c_double)]
    lib.oracle_trace_sample.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_uint64, C.c_uint64,
                                        P(C.c_double)]
    lib.oracle_hg.argtypes = [C.c_double, C.c_int, P(C.c_double), P(C.c_double), C.c_double, C.c_double,
                              P(C.c_double)]
    lib.oracle_filter_axis.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, P(C.c_int32), P(C.c_float), C.c_int]
    _lib = lib
    return lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _d(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


COUNTER_NAMES = ("closest", "shadow", "node_visits", "tri_tests", "hits")


class OracleScene:
    """A scene restated on the CPU in float (precision=32) or double (64)."""

    def __init__(self, desc, precision: int = 32, eps: float = 3e-4):
        self.lib = load()
        self.h = self.lib.oracle_scene_create(C.byref(desc), precision, eps)
        if not self.h:
            raise RuntimeError("oracle scene: " + self.lib.oracle_last_error().decode())

    def close(self):
        if self.h:
            self.lib.oracle_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def trace_rays(self, o, d, tmin, tmax, any_hit=False):
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
        cnt = np.zeros(5, np.uint64)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        rc = self.lib.oracle_trace_rays(self.h, _f(o), _f(d), _f(tmin), _f(tmax), n, 1 if any_hit else 0, ip(ent),
                                        ip(prim), _f(t), _f(u), _f(v), cnt.ctypes.data_as(C.POINTER(C.c_uint64)))
        if rc:
            raise RuntimeError(self.lib.oracle_last_error().decode())
        return {"entity": ent, "prim": prim, "t": t, "u": u, "v": v, "counters": dict(zip(COUNTER_NAMES, cnt.tolist()))}

    def medium(self, mi: int, mode: str, a, b, n: int, seed: int = 1, indirect: bool = False) -> np.ndarray:
        """n independent trials of Medium::tr ('tr'), ::ab ('ab') or ::sample_scattering
        ('scatter': columns happened, pos[3], wr[3], throughput[3], g) along a -> b."""
        m = {"tr": 0, "ab": 1, "scatter": 2}[mode]
        width = 11 if m == 2 else 3
        out = np.zeros((n, width), np.float64)
        a = np.ascontiguousarray(a, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        if self.lib.oracle_medium(self.h, mi, m, _d(a), _d(b), int(indirect), seed, n, _d(out)) != 0:
            raise RuntimeError(self.lib.oracle_last_error().decode())
        return out

    def tex3d(self, mi: int, which: str, uvw) -> np.ndarray:
        uvw = np.ascontiguousarray(np.atleast_2d(uvw), np.float64)
        out = np.zeros((len(uvw), 3), np.float64)
        w = {"density": 0, "albedo": 1, "g": 2}[which]
        if self.lib.oracle_tex3d(self.h, mi, w, _d(uvw), len(uvw), _d(out)) != 0:
            raise RuntimeError(self.lib.oracle_last_error().decode())
        return out

    def trace_sample(self, params, qx: int, qy: int, key: int, sample: int) -> np.ndarray:
        """One camera sample: value[3], albedo[3], normal[3], denoise, film x, film y."""
        out = np.zeros(12, np.float64)
        if self.lib.oracle_trace_sample(self.h, C.byref(params), qx, qy, key, sample, _d(out)) != 0:
            raise RuntimeError(self.lib.oracle_last_error().decode())
        return out

    def env(self, mode: str, x) -> np.ndarray:
        """Environment light: mode "radiance" / "pdf" on directions (n, 3), or
        "sample" on uniforms (n, 3) -> (n, 4) = direction, pdf."""
        x = np.ascontiguousarray(np.atleast_2d(x), np.float64)
        out = np.zeros((len(x), 4), np.float64)
        m = {"radiance": 0, "pdf": 1, "sample": 2}[mode]
        if self.lib.oracle_env(self.h, m, _d(x), len(x), _d(out)) != 0:
            raise RuntimeError(self.lib.oracle_last_error().decode())
        return out[:, :3] if m == 0 else (out[:, 0] if m == 1 else out)

    def render(self, params, threads: int | None = None, rows: tuple | None = None, aovs: bool = True):
        W, H = params.width, params.height
        rgb = np.zeros((H, W, 3), np.float64)
        alb = np.zeros((H, W, 3), np.float64) if aovs else None
        nor = np.zeros((H, W, 3), np.float64) if aovs else None
        den = np.zeros((H, W), np.float64) if aovs else None
        cnt = np.zeros(5, np.uint64)
        secs = C.c_double()
        r0, r1 = rows if rows else (0, -1)
        th = threads if threads else (os.cpu_count() or 1)
        rc = self.lib.oracle_render(self.h, C.byref(params), th, _d(rgb), _d(alb), _d(nor), _d(den),
                                    cnt.ctypes.data_as(C.POINTER(C.c_uint64)), r0, r1, C.byref(secs))
        if rc:
            raise RuntimeError(self.lib.oracle_last_error().decode())
        return {"image": rgb, "albedo": alb, "normal": nor, "denoise": den, "seconds": secs.value,
                "counters": dict(zip(COUNTER_NAMES, cnt.tolist())), "threads": th}


def triangle_soup_closest(tris, o, d, tmin, tmax, watertight: bool):
    """Brute-force closest hit over triangles (n, 3, 3) with the watertight test
    or Atrc's Moller-Trumbore test.  Returns (prim, t, u, v); prim = -1 on a miss."""
    lib = load()
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    o = np.ascontiguousarray(o, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
    n = len(o)
    tmin = np.ascontiguousarray(np.broadcast_to(np.asarray(tmin, np.float32), (n,)))
    tmax = np.ascontiguousarray(np.broadcast_to(np.asarray(tmax, np.float32), (n,)))
    prim = np.empty(n, np.int32)
    t, u, v = (np.empty(n, np.float32) for _ in range(3))
    lib.oracle_triangle_soup_closest(_f(tris), len(tris), _f(o), _f(d), _f(tmin), _f(tmax), n,
                                     1 if watertight else 0, prim.ctypes.data_as(C.POINTER(C.c_int32)),
                                     _f(t), _f(u), _f(v))
    return prim, t, u, v


def pcg_stream(seed: int, pixel: int, sample: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.float32)
    load().oracle_pcg_stream(seed, pixel, sample, n, _f(out))
    return out


def pcg_raw(initstate: int, initseq: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    load().oracle_pcg_raw(initstate, initseq, n, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


def distribution(kind: int, u: float, v: float, w: float = 0.0) -> np.ndarray:
    out = np.zeros(9, np.float64)
    load().oracle_distribution(kind, u, v, w, _d(out))
    return out


def bsdf(material_desc, mode: int, wi=(0, 0, 1), wo=(0, 0, 1), u=0.5, v=0.5, w=0.5, precision=64, geo_z=None):
    out = np.zeros(8, np.float64)
    wi = np.asarray(wi, np.float64)
    wo = np.asarray(wo, np.float64)
    gz = None if geo_z is None else np.asarray(geo_z, np.float64)
    load().oracle_bsdf(C.byref(material_desc), precision, _d(gz), _d(wi), _d(wo), u, v, w, mode, _d(out))
    return out


def hg_eval(g: float, wi, wo) -> float:
    out = np.zeros(1, np.float64)
    load().oracle_hg(g, 0, _d(np.asarray(wi, np.float64)), _d(np.asarray(wo, np.float64)), 0.0, 0.0, _d(out))
    return float(out[0])


def hg_sample(g: float, wo, u: float, v: float):
    out = np.zeros(4, np.float64)
    z = np.zeros(3, np.float64)
    load().oracle_hg(g, 1, _d(z), _d(np.asarray(wo, np.float64)), u, v, _d(out))
    return out[:3], float(out[3])


def filter_axis(lo: int, hi: int, r: float, s: float) -> list:
    """Pixels of one axis that Atrc's apply_image_filter reaches from a sample at s
    in a film grid with pixel range [lo, hi], with the distances it passes."""
    pix = np.zeros(256, np.int32)
    rel = np.zeros(256, np.float32)
    n = load().oracle_filter_axis(lo, hi, r, s, pix.ctypes.data_as(C.POINTER(C.c_int32)), _f(rel), 256)
    return [(int(pix[i]), float(rel[i])) for i in range(n)]
