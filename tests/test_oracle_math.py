"""Pins of the CPU oracle that need no GPU: the random-number generator, the
sampling distributions, the ray/triangle test and the BVH, each against a
closed form, a published known answer or brute force.

Atrc ships no tests or golden vectors (its only fixtures are gallery PNGs
without scene files), so these are the pins.  Functions that Atrc takes from
its absent agz-utils submodule (distributions, frames, alias table) follow the
published method named in oracle/atrc_oracle.hpp and are pinned here against
that method's closed form.
"""
import math

import numpy as np
import pytest

from atrc_amd import capi, scenes
from oracle import py_oracle as po


# --------------------------------------------------------------------------- RNG

def test_pcg32_known_answer():
    # pcg32_srandom_r(42, 54): the first outputs printed by the PCG reference demo
    # (pcg-c, pcg32-demo, "Round 1"); Atrc's NativeSampler is pcg32 (core/sampler.h:28)
    out = po.pcg_raw(42, 54, 6)
    assert [hex(x) for x in out] == ["0xa15c02b7", "0x7b47f409", "0xba1d3330", "0x83d2f293", "0xbfa4784b",
                                     "0xcbed606e"]


def test_uniform_stream_range_and_moments():
    u = np.concatenate([po.pcg_stream(42, p, s, 64) for p in range(64) for s in range(4)])
    assert u.min() >= 0.0 and u.max() < 1.0
    # floats are k * 2^-24 exactly
    assert np.all((u * 2 ** 24) == np.floor(u * 2 ** 24))
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.005
    # distinct (pixel, sample) streams differ
    assert not np.array_equal(po.pcg_stream(42, 0, 0, 16), po.pcg_stream(42, 0, 1, 16))
    assert not np.array_equal(po.pcg_stream(42, 0, 0, 16), po.pcg_stream(42, 1, 0, 16))


# --------------------------------------------------------------------------- distributions

def test_concentric_disk_and_cosine_hemisphere():
    rng = np.random.default_rng(1)
    pts = np.array([po.distribution(0, u, v)[:2] for u, v in rng.uniform(size=(4000, 2))])
    r2 = (pts ** 2).sum(1)
    assert r2.max() <= 1 + 1e-12
    # uniform on the disk: E[r^2] = 1/2, and quadrant balance
    assert abs(r2.mean() - 0.5) < 0.02
    # corners of the square map to the rim, the centre to the origin
    assert np.allclose(po.distribution(0, 0.5, 0.5)[:2], 0)
    assert abs(np.hypot(*po.distribution(0, 1.0, 0.5)[:2]) - 1) < 1e-12
    # cosine hemisphere: pdf = z / pi, E[z] = 2/3 under cos-weighting
    s = np.array([po.distribution(1, u, v) for u, v in rng.uniform(size=(4000, 2))])
    assert np.allclose(np.linalg.norm(s[:, :3], axis=1), 1, atol=1e-12)
    assert np.allclose(s[:, 3], s[:, 2] / math.pi)
    assert abs(s[:, 2].mean() - 2 / 3) < 0.01


def test_sphere_cone_triangle_maps():
    assert np.allclose(po.distribution(2, 0.0, 0.3)[:3], [0, 0, 1])     # z = 1 - 2u
    assert np.allclose(po.distribution(2, 1.0, 0.3)[:3], [0, 0, -1])
    c = po.distribution(3, 1.0, 0.0, 0.5)                                # u = 1 -> cos = cos_max
    assert abs(c[2] - 0.5) < 1e-12 and abs(c[3] - 1 / (2 * math.pi * 0.5)) < 1e-12
    b = po.distribution(4, 0.25, 0.5)                                    # (1 - sqrt u, v sqrt u)
    assert np.allclose(b[:2], [0.5, 0.25])
    rng = np.random.default_rng(2)
    bs = np.array([po.distribution(4, u, v)[:2] for u, v in rng.uniform(size=(6000, 2))])
    assert (bs >= 0).all() and (bs.sum(1) <= 1 + 1e-12).all()
    assert np.allclose(bs.mean(0), [1 / 3, 1 / 3], atol=0.01)           # uniform: centroid


def test_orthonormal_basis():
    rng = np.random.default_rng(3)
    for z in rng.normal(size=(200, 3)).tolist() + [[0, 0, 1], [0, 0, -1], [1, 0, 0]]:
        f = po.distribution(5, *z).reshape(3, 3)
        zn = np.array(z) / np.linalg.norm(z)
        assert np.allclose(f[2], zn)
        assert np.allclose(f @ f.T, np.eye(3), atol=1e-12)
        assert np.allclose(np.cross(f[0], f[1]), f[2], atol=1e-12)      # right-handed


# --------------------------------------------------------------------------- intersection

def _tri_scene(tris, aggregate=capi.AGGREGATE_NATIVE):
    s = scenes.Scene(aggregate=aggregate)
    m = s.add_material(scenes.Material())
    s.add(scenes.Entity(capi.GEOM_TRIANGLE_BVH, m, tris=tris))
    return s


def _mesh(verts):
    out = np.zeros((len(verts), 8), np.float32)
    out[:, :3] = verts
    out[:, 5] = 1.0
    return out


def test_mt_known_answer_dyadic():
    """Dyadic triangle and ray: Möller–Trumbore is exact, so (t, alpha, beta) are
    the exact rationals (triangle_aux.h:38-69: alpha weights B-A, beta C-A)."""
    s = _tri_scene(_mesh([[0, 0, 0], [1, 0, 0], [0, 1, 0]]))
    d, k = s.to_desc()
    o = po.OracleScene(d, 32)
    r = o.trace_rays(np.array([[0.25, 0.5, 2.0]]), np.array([[0, 0, -1.0]]), np.zeros(1), np.full(1, np.inf))
    assert r["entity"][0] == 0 and r["t"][0] == 2.0 and r["u"][0] == 0.25 and r["v"][0] == 0.5
    # inclusive bounds: t exactly t_max still hits (common.h:320-323)
    r = o.trace_rays(np.array([[0.25, 0.5, 2.0]]), np.array([[0, 0, -1.0]]), np.zeros(1), np.full(1, 2.0))
    assert r["entity"][0] == 0
    r = o.trace_rays(np.array([[0.25, 0.5, 2.0]]), np.array([[0, 0, -1.0]]), np.zeros(1), np.full(1, 1.9999))
    assert r["entity"][0] == -1
    # exactly on the edge alpha + beta = 1 counts as a hit
    r = o.trace_rays(np.array([[0.5, 0.5, 1.0]]), np.array([[0, 0, -1.0]]), np.zeros(1), np.full(1, np.inf))
    assert r["entity"][0] == 0 and r["u"][0] + r["v"][0] == 1.0
    # ray in the triangle's plane: det == 0, no hit
    r = o.trace_rays(np.array([[-1, 0.25, 0.0]]), np.array([[1.0, 0, 0]]), np.zeros(1), np.full(1, np.inf))
    assert r["entity"][0] == -1


def test_tie_rule_last_primitive_and_last_entity_win():
    # two coincident triangles in one leaf: the later one in BVH order wins an equal-t tie
    tri = [[0, 0, 0], [1, 0, 0], [0, 1, 0]]
    s = _tri_scene(_mesh(tri + tri))
    d, k = s.to_desc()
    r = po.OracleScene(d, 32).trace_rays(np.array([[0.2, 0.2, 1.0]]), np.array([[0, 0, -1.0]]), np.zeros(1),
                                         np.full(1, np.inf))
    assert r["prim"][0] == 1
    # two coincident quads under the native aggregate: the later entity wins (native_aggregate.cpp:31-45)
    s = scenes.Scene()
    m = s.add_material(scenes.Material())
    q = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0))
    s.add(scenes.Entity(capi.GEOM_QUAD, m, points=q))
    s.add(scenes.Entity(capi.GEOM_QUAD, m, points=q))
    d, k = s.to_desc()
    r = po.OracleScene(d, 32).trace_rays(np.array([[0.7, 0.2, 1.0]]), np.array([[0, 0, -1.0]]), np.zeros(1),
                                         np.full(1, np.inf))
    assert r["entity"][0] == 1 and r["prim"][0] == 0


def _mt_numpy(o, d, A, BA, CA):
    """Möller–Trumbore in float32 numpy with Atrc's operation order (no FMA)."""
    f = np.float32
    s1 = np.cross(d, CA).astype(f)
    div = (s1[..., 0] * BA[..., 0] + s1[..., 1] * BA[..., 1]) + s1[..., 2] * BA[..., 2]
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = f(1) / div
        oa = (o - A).astype(f)
        alpha = ((oa[..., 0] * s1[..., 0] + oa[..., 1] * s1[..., 1]) + oa[..., 2] * s1[..., 2]) * inv
        s2 = np.stack([oa[..., 1] * BA[..., 2] - oa[..., 2] * BA[..., 1], oa[..., 2] * BA[..., 0] - oa[..., 0] * BA[..., 2],
                       oa[..., 0] * BA[..., 1] - oa[..., 1] * BA[..., 0]], -1).astype(f)
        beta = ((d[..., 0] * s2[..., 0] + d[..., 1] * s2[..., 1]) + d[..., 2] * s2[..., 2]) * inv
        t = ((CA[..., 0] * s2[..., 0] + CA[..., 1] * s2[..., 1]) + CA[..., 2] * s2[..., 2]) * inv
    ok = (div != 0) & (alpha >= 0) & (beta >= 0) & (alpha + beta <= 1) & (t >= 0)
    return np.where(ok, t, np.inf)


def test_bvh_closest_equals_brute_force():
    """The oracle's BVH traversal returns the nearest hit of a brute-force scan."""
    rng = np.random.default_rng(5)
    n = 1500
    c = rng.uniform(-5, 5, (n, 1, 3))
    v = (c + rng.normal(scale=0.6, size=(n, 3, 3))).astype(np.float32).reshape(-1, 3)
    s = _tri_scene(_mesh(v))
    d, k = s.to_desc()
    o, dd = scenes.random_rays(((-6, -6, -6), (6, 6, 6)), 3000, seed=6)
    r = po.OracleScene(d, 32).trace_rays(o, dd, np.zeros(len(o)), np.full(len(o), np.inf))
    tri = v.reshape(-1, 3, 3)
    A, BA, CA = tri[:, 0], (tri[:, 1] - tri[:, 0]).astype(np.float32), (tri[:, 2] - tri[:, 0]).astype(np.float32)
    tt = _mt_numpy(o[:, None, :], dd[:, None, :], A[None], BA[None], CA[None])
    best = tt.min(axis=1)
    hit = np.isfinite(best)
    assert (hit == (r["entity"] >= 0)).all()
    np.testing.assert_array_equal(r["t"][hit], best[hit])
    # the reported primitive is one that attains the minimum
    assert (tt[np.arange(len(o))[hit], r["prim"][hit]] == best[hit]).all()


# --------------------------------------------------------------------------- BVH build

def _restated_build(tris):
    """Atrc's split rule (triangle_bvh.cpp:122-223) restated in numpy float32 for
    small inputs: returns the DFS node list [(lo, hi, leaf_ids or None)]."""
    tris = tris.reshape(-1, 3, 3).astype(np.float32)
    cen = ((tris[:, 0] + tris[:, 1]) + tris[:, 2]) / np.float32(3)
    order = list(range(len(tris)))

    def build(lo_i, hi_i, depth):
        ids = order[lo_i:hi_i]
        pts = tris[ids].reshape(-1, 3)
        box = (pts.min(0), pts.max(0))
        if len(ids) <= 5:
            return (box, list(ids), None, None)
        cb_lo, cb_hi = cen[ids].min(0), cen[ids].max(0)
        e = cb_hi - cb_lo
        axis = (0 if e[0] > e[2] else 2) if e[0] > e[1] else (1 if e[1] > e[2] else 2)
        split = np.float32(0.5) * (cb_hi[axis] + cb_lo[axis])
        mid = lo_i
        for i in range(lo_i, hi_i):
            if cen[order[i]][axis] < split:
                order[i], order[mid] = order[mid], order[i]
                mid += 1
        if mid == lo_i or mid == hi_i:
            mid = lo_i + len(ids) // 2
        return (box, None, build(lo_i, mid, depth + 1), build(mid, hi_i, depth + 1))

    root = build(0, len(tris), 0)
    flat = []

    def dfs(node):
        flat.append(node)
        if node[1] is None:
            dfs(node[2])
            dfs(node[3])

    dfs(root)
    return flat


@pytest.mark.parametrize("n,seed", [(8, 1), (17, 2), (40, 3), (64, 4)])
def test_product_bvh_matches_restated_split_rule(n, seed):
    """libatrc_gpu.so's Atrc BVH builder (used for the GPU traversal) against an
    independent numpy restatement of the split rule; bounds, order and leaf
    contents must be identical."""
    rng = np.random.default_rng(seed)
    v = rng.uniform(-3, 3, (n * 3, 3)).astype(np.float32)
    nodes, prims, depth = capi.debug_triangle_bvh(_mesh(v))
    ref = _restated_build(v)
    assert len(nodes) == len(ref)
    ids = prims[:, 3].view(np.uint32)
    for k, (box, leaf, _, _) in enumerate(ref):
        np.testing.assert_array_equal(nodes[k, 0:3], box[0])
        np.testing.assert_array_equal(nodes[k, 4:7], box[1])
        start, end = nodes[k, 3:4].view(np.uint32)[0], nodes[k, 7:8].view(np.uint32)[0]
        if leaf is None:
            assert start == 0xFFFFFFFF
        else:
            assert list(ids[start:end]) == leaf


def test_product_bvh_invariants_large():
    s, _, _ = scenes.procedural_spheres(n_spheres=3, subdiv=4, seed=9)
    tris = s.entities[0].tris
    nodes, prims, depth = capi.debug_triangle_bvh(tris)
    n = len(tris) // 3
    start = nodes[:, 3].view(np.uint32)
    end = nodes[:, 7].view(np.uint32)
    leaf = start != 0xFFFFFFFF
    assert ((end[leaf] - start[leaf]) <= 5).all() and ((end[leaf] - start[leaf]) >= 1).all()
    # every triangle exactly once
    ids = prims[:, 3].view(np.uint32)
    assert sorted(ids.tolist()) == list(range(n))
    assert (np.diff(start[leaf].astype(np.int64)) > 0).all()
    # interior: left child follows, bounds contain both children
    for i in np.nonzero(~leaf)[0]:
        for ch in (i + 1, end[i]):
            assert (nodes[ch, 0:3] >= nodes[i, 0:3]).all() and (nodes[ch, 4:7] <= nodes[i, 4:7]).all()
    assert depth < 64
