"""The benchmark workloads themselves, on the GPU, against the oracle.

* 1M-triangle scene (200 displaced icospheres at subdivision 4, ground quad):
  300k seeded rays (random rays in the scene box and rays from the camera)
  through Atrc's own BVHs (AGX_ACCEL_ATRC) equal the oracle bit for bit:
  entity and primitive ids, t, u, v, any-hit results, and the total node
  visits and triangle tests (triangle_bvh.cpp:363-488, entity_bvh.cpp).  The
  production BVH4 gives the same hits except where two primitives meet a ray
  at exactly the same t.
* A full 1920x1080 frame of that scene (BVH4, depth 8): finite, bit-identical
  when rendered twice, and bit-identical when the context's pass capacity is
  cut to 2^20 paths (a sample layer of 2.07M paths then spans several passes
  and the film takes partial layers).
* The door-gap room (500k clutter triangles) and the volumetric room (the same
  room filled with a 256^3 fBm heterogeneous medium and a homogeneous slab):
  path for path at 1 spp against the FP32 oracle on a band of rows of the
  GPU's full frame.
"""
import numpy as np
import pytest

from atrc_amd import capi, scenes
from oracle import py_oracle as po

from conftest import log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mesh1m():
    s, cam, tr = scenes.procedural_spheres(n_spheres=200, subdiv=4, seed=1)
    assert s.triangle_count() == 1_024_000
    return s, cam, tr


def _camera_rays(cam, n, seed):
    rng = np.random.default_rng(seed)
    o = np.tile(np.asarray(cam["pos"], np.float32), (n, 1))
    tgt = rng.uniform(0.0, 50.0, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o.astype(np.float32), d.astype(np.float32)


@pytest.mark.parametrize("accel", [capi.ACCEL_ATRC, capi.ACCEL_BVH4])
def test_mesh1m_rays(gpu_ctx, mesh1m, accel):
    s, cam, tr = mesh1m
    s.accel = accel
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    ref = po.OracleScene(desc, 32)
    o1, d1 = scenes.random_rays(((0, 0, 0), (50, 50, 50)), 200_000, seed=41)
    o2, d2 = _camera_rays(cam, 100_000, seed=42)
    o, d = np.concatenate([o1, o2]), np.concatenate([d1, d2])
    n = len(o)
    tmin, tmax = np.zeros(n, np.float32), np.full(n, np.inf, np.float32)
    log(f"[workloads] mesh1m rays, accel {accel}: {n} rays")
    g = gpu_ctx.trace_rays(o, d, tmin, tmax)
    st = gpu_ctx.stats()
    r = ref.trace_rays(o, d, tmin, tmax)
    hit = r["entity"] >= 0
    same = (g["entity"] == r["entity"]) & (g["prim"] == r["prim"])
    log(f"  hits {hit.mean():.3f}; differing winners {int((~same).sum())}")
    if accel == capi.ACCEL_ATRC:
        assert same.all()
        assert st["closest_node_visits"] == r["counters"]["node_visits"]
        assert st["closest_tri_tests"] == r["counters"]["tri_tests"]
    else:
        assert (~same).mean() < 1e-3
        np.testing.assert_array_equal(g["t"][~same].view(np.uint32), r["t"][~same].view(np.uint32))
        hit &= same
    for k in ("t", "u", "v"):
        np.testing.assert_array_equal(g[k][hit].view(np.uint32), r[k][hit].view(np.uint32))
    ga = gpu_ctx.trace_rays(o, d, tmin, tmax, any_hit=True)
    ra = ref.trace_rays(o, d, tmin, tmax, any_hit=True)
    np.testing.assert_array_equal(ga["entity"], ra["entity"])


def test_mesh1m_full_frame(gpu_ctx, mesh1m, monkeypatch):
    s, cam, tr = mesh1m
    s.accel = capi.ACCEL_BVH4
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    p = scenes.render_params(1920, 1080, 4, cam, seed=61, **tr)
    log("[workloads] mesh1m 1920x1080 4 spp: twice, then with passes of 2^20 paths")
    a = gpu_ctx.render(p)
    b = gpu_ctx.render(p)
    for k in ("image", "albedo", "normal", "denoise"):
        assert np.isfinite(a[k]).all(), k
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    assert a["image"].mean() > 0.05
    monkeypatch.setenv("AGX_MAX_PATHS", str(1 << 20))
    with capi.Context(0) as small:
        small.upload_scene(desc)
        c = small.render(p)
    for k in ("image", "albedo", "normal", "denoise"):
        assert np.array_equal(a[k].view(np.uint32), c[k].view(np.uint32)), k


def _crop_parity(gpu_ctx, s, cam, tr, W, H, rows, label):
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    p = scenes.render_params(W, H, 1, cam, seed=71, **tr)
    log(f"[workloads] {label} {W}x{H} 1 spp, oracle rows {rows[0]}..{rows[1]}")
    g = gpu_ctx.render(p)
    r = po.OracleScene(desc, 32).render(p, threads=8, rows=rows)
    sl = slice(*rows)
    gi, ri = g["image"][sl].astype(np.float64), r["image"][sl]
    assert np.isfinite(g["image"]).all()
    rel = np.abs(gi - ri).max(axis=2) / (np.abs(ri).max(axis=2) + 1e-2)
    frac = float((rel <= 1e-3).mean())
    lit = float((ri.max(axis=2) > 0).mean())
    log(f"  per-path agreement {frac:.5f} (lit pixels {lit:.3f})")
    assert frac >= 0.98
    np.testing.assert_allclose(g["albedo"][sl], r["albedo"][sl], atol=1e-4)
    np.testing.assert_allclose(g["normal"][sl], r["normal"][sl], atol=1e-4)


@pytest.mark.parametrize("accel", [capi.ACCEL_ATRC, capi.ACCEL_BVH4])
def test_door_gap_room_path_parity(gpu_ctx, accel):
    s, cam, tr = scenes.guided_room(n_clutter_tris=500_000, seed=7)
    s.accel = accel
    _crop_parity(gpu_ctx, s, cam, tr, 512, 512, (160, 288), f"door-gap room (accel {accel})")


def test_volumetric_room_path_parity(gpu_ctx):
    s, cam, tr = scenes.volumetric_room(n_clutter_tris=500_000, grid_n=256)
    s.accel = capi.ACCEL_BVH4
    _crop_parity(gpu_ctx, s, cam, tr, 512, 512, (192, 256), "volumetric room (256^3 grid)")
