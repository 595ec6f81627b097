"""Ray-query parity: the gfx950 traversal against the CPU restatement of Atrc's
scene queries (oracle, FP32), ray for ray.

Bar: bit-exact.  Entity id, primitive id (original triangle index / quad half),
hit distance t and barycentrics (u, v) must be identical, and any-hit results
identical, because the GPU traverses the same Atrc BVHs with the same FP32
operation order (Möller–Trumbore, inclusive t bounds, last-tested wins ties:
src/tracer/include/agz/tracer/utility/triangle_aux.h:38-69,
src/tracer/src/core/geometry/triangle_bvh.cpp:403-488).
"""
import numpy as np
import pytest

from atrc_amd import capi, scenes
from oracle import py_oracle as po

from conftest import log

pytestmark = pytest.mark.gpu


def _compare(gpu_ctx, scene, o, d, tmin, tmax, label):
    desc, keep = scene.to_desc()
    gpu_ctx.upload_scene(desc)
    ref = po.OracleScene(desc, 32)
    log(f"[trace parity] {label}: {len(o)} rays")
    g = gpu_ctx.trace_rays(o, d, tmin, tmax)
    r = ref.trace_rays(o, d, tmin, tmax)
    np.testing.assert_array_equal(g["entity"], r["entity"])
    np.testing.assert_array_equal(g["prim"], r["prim"])
    hit = r["entity"] >= 0
    # bit-exact FP32 distances and barycentrics
    np.testing.assert_array_equal(g["t"][hit].view(np.uint32), r["t"][hit].view(np.uint32))
    np.testing.assert_array_equal(g["u"][hit].view(np.uint32), r["u"][hit].view(np.uint32))
    np.testing.assert_array_equal(g["v"][hit].view(np.uint32), r["v"][hit].view(np.uint32))
    st = gpu_ctx.stats()
    # same BVH, same order: node visits and triangle tests agree with the oracle
    assert st["closest_node_visits"] == r["counters"]["node_visits"]
    assert st["closest_tri_tests"] == r["counters"]["tri_tests"]
    ga = gpu_ctx.trace_rays(o, d, tmin, tmax, any_hit=True)
    ra = ref.trace_rays(o, d, tmin, tmax, any_hit=True)
    np.testing.assert_array_equal(ga["entity"], ra["entity"])
    return hit.mean()


def test_cornell_rays(gpu_ctx):
    s, _, _ = scenes.cornell_box()
    o, d = scenes.random_rays(((0.05, 0.05, 0.05), (5.5, 5.5, 5.5)), 200_000, seed=7)
    n = len(o)
    frac = _compare(gpu_ctx, s, o, d, np.zeros(n, np.float32), np.full(n, np.inf, np.float32), "cornell")
    assert frac > 0.8  # the box is open at the front (y = 0)


def test_cornell_native_aggregate_and_bounded_rays(gpu_ctx):
    s, _, _ = scenes.cornell_box()
    s.aggregate = capi.AGGREGATE_NATIVE
    o, d = scenes.random_rays(((0.05, 0.05, 0.05), (5.5, 5.5, 5.5)), 100_000, seed=8)
    rng = np.random.default_rng(9)
    tmin = rng.uniform(0, 0.5, len(o)).astype(np.float32)
    tmax = rng.uniform(0.5, 6.0, len(o)).astype(np.float32)
    _compare(gpu_ctx, s, o, d, tmin, tmax, "cornell/native/bounded")


def test_procedural_mesh_rays(gpu_ctx):
    s, _, _ = scenes.procedural_spheres(n_spheres=24, subdiv=3, seed=3, extent=20.0)
    o, d = scenes.random_rays(((-2, -2, -2), (22, 22, 22)), 200_000, seed=11)
    n = len(o)
    frac = _compare(gpu_ctx, s, o, d, np.zeros(n, np.float32), np.full(n, np.inf, np.float32), "spheres24")
    assert frac > 0.3


def test_analytic_shapes_rays(gpu_ctx):
    """Transformed triangle, sphere and disk entities (local-space intersection)."""
    s = scenes.Scene(aggregate=capi.AGGREGATE_BVH)
    m = s.add_material(scenes.Material(color=(0.5, 0.5, 0.5)))
    rot = np.array([[0.6, -0.8, 0.0], [0.8, 0.6, 0.0], [0.0, 0.0, 1.0]])
    for k in range(6):
        M = np.hstack([rot * (0.5 + 0.25 * k), np.array([[k * 2.0], [1.0], [0.5 * k]])])
        s.add(scenes.Entity(capi.GEOM_SPHERE, m, radius=1.0, local_to_world=tuple(M.reshape(-1))))
        s.add(scenes.Entity(capi.GEOM_DISK, m, radius=0.8, local_to_world=tuple((M + np.array(
            [[0, 0, 0, 0], [0, 0, 0, 3.0], [0, 0, 0, 0]])).reshape(-1))))
        s.add(scenes.Entity(capi.GEOM_TRIANGLE, m, points=((0, 0, 0), (1, 0, 0), (0, 1, 0)),
                            local_to_world=tuple((M + np.array([[0, 0, 0, 0], [0, 0, 0, -3.0], [0, 0, 0, 0]])).reshape(-1))))
    o, d = scenes.random_rays(((-2, -4, -2), (13, 6, 5)), 100_000, seed=12)
    n = len(o)
    _compare(gpu_ctx, s, o, d, np.zeros(n, np.float32), np.full(n, np.inf, np.float32), "analytic")
