"""AGX_ACCEL_BVH4_WATERTIGHT on the gfx950 path against the oracle's
watertight test (tests/test_oracle_watertight.py pins it):

* every GPU hit is the oracle's watertight hit of the same triangle, bit for
  bit (t, u, v), at the smallest t over the whole mesh (brute force);
* rays from inside a closed mesh aimed exactly at its vertices and edges never
  escape (the same rays leak through Atrc's Moller-Trumbore test);
* a render in this mode matches the oracle path for path wherever no ray meets
  an edge.
"""
import numpy as np
import pytest

from atrc_amd import capi, scenes
from oracle import py_oracle as po

from conftest import log
from test_gpu_render_parity import _path_parity
from test_oracle_watertight import _closed_mesh

pytestmark = pytest.mark.gpu


def _mesh_scene(accel):
    pos, faces, tris = _closed_mesh(subdiv=3)
    s = scenes.Scene(aggregate=capi.AGGREGATE_BVH)
    m = s.add_material(scenes.Material(color=(0.5, 0.5, 0.5)))
    s.add(scenes.Entity(capi.GEOM_TRIANGLE_BVH, m,
                        tris=scenes.mesh_from_indexed(pos, scenes.vertex_normals(pos, faces), faces)))
    s.accel = accel
    return s, pos, faces, tris


def _edge_rays(pos, faces, seed, n_origins=6):
    edges = set()
    for f in faces:
        for a, b in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0])):
            edges.add((min(a, b), max(a, b)))
    edges = np.array(sorted(edges))
    rng = np.random.default_rng(seed)
    O, D = [], []
    for _ in range(n_origins):
        o = (np.float32([0.3, -0.2, 0.9]) + rng.uniform(-0.5, 0.5, 3)).astype(np.float32)
        ts = rng.random(3).astype(np.float32)
        targets = np.concatenate([pos] + [(pos[edges[:, 0]] * (1 - t) + pos[edges[:, 1]] * t).astype(np.float32)
                                          for t in ts])
        O.append(np.broadcast_to(o, targets.shape))
        D.append((targets - o).astype(np.float32))
    return np.concatenate(O).astype(np.float32), np.concatenate(D).astype(np.float32)


def test_watertight_hits_match_the_oracle(gpu_ctx):
    s, pos, faces, tris = _mesh_scene(capi.ACCEL_BVH4_WATERTIGHT)
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    o, d = _edge_rays(pos, faces, seed=11)
    rng = np.random.default_rng(12)
    # plus generic rays from outside and inside
    ro = rng.uniform(-3, 3, (20000, 3)).astype(np.float32)
    rd = rng.normal(size=(20000, 3)).astype(np.float32)
    o = np.concatenate([o, ro])
    d = np.concatenate([d, rd])
    n = len(o)
    log(f"[watertight] {n} rays ({n - 20000} aimed at vertices / edges from inside) vs oracle brute force")
    g = gpu_ctx.trace_rays(o, d, np.zeros(n, np.float32), np.full(n, np.inf, np.float32))
    prim, t, u, v = po.triangle_soup_closest(tris, o, d, 0, np.inf, watertight=True)
    np.testing.assert_array_equal(g["prim"] >= 0, prim >= 0)
    hit = prim >= 0
    # the smallest t is unique as a value even where two triangles tie at it
    np.testing.assert_array_equal(g["t"][hit], t[hit])
    # each GPU hit is the oracle's test of that same triangle
    idx = np.nonzero(hit)[0]
    p1, t1, u1, v1 = (np.empty(len(idx), x.dtype) for x in (prim, t, u, v))
    for k, i in enumerate(idx):
        r = po.triangle_soup_closest(tris[g["prim"][i]:g["prim"][i] + 1], o[i:i + 1], d[i:i + 1], 0, np.inf,
                                     watertight=True)
        p1[k], t1[k], u1[k], v1[k] = (x[0] for x in r)
    assert (p1 == 0).all()
    np.testing.assert_array_equal(t1, g["t"][idx])
    np.testing.assert_array_equal(u1, g["u"][idx])
    np.testing.assert_array_equal(v1, g["v"][idx])
    same_prim = (g["prim"] == prim)[hit].mean()
    log(f"  all hits bit-exact; same triangle as the brute force on {same_prim:.4%} (the rest tie at equal t)")


def test_closed_mesh_does_not_leak(gpu_ctx):
    pos, faces, tris = _closed_mesh(subdiv=3)
    o, d = _edge_rays(pos, faces, seed=21, n_origins=10)
    n = len(o)
    leaks = {}
    for accel in (capi.ACCEL_BVH4_WATERTIGHT, capi.ACCEL_BVH4):
        s, _, _, _ = _mesh_scene(accel)
        desc, keep = s.to_desc()
        gpu_ctx.upload_scene(desc)
        g = gpu_ctx.trace_rays(o, d, np.zeros(n, np.float32), np.full(n, np.inf, np.float32))
        leaks[accel] = int((g["prim"] < 0).sum())
        a = gpu_ctx.trace_rays(o, d, np.zeros(n, np.float32), np.full(n, np.inf, np.float32), any_hit=True)
        leaks[(accel, "any")] = int((a["entity"] < 0).sum())
    log(f"[watertight] {n} edge / vertex rays from inside: escaped watertight {leaks[capi.ACCEL_BVH4_WATERTIGHT]}, "
        f"Moller-Trumbore {leaks[capi.ACCEL_BVH4]}")
    assert leaks[capi.ACCEL_BVH4_WATERTIGHT] == 0
    assert leaks[(capi.ACCEL_BVH4_WATERTIGHT, "any")] == 0
    assert leaks[capi.ACCEL_BVH4] > 0


def test_watertight_render_path_parity(gpu_ctx):
    s, cam, tr = scenes.cornell_box()
    s.accel = capi.ACCEL_BVH4_WATERTIGHT
    _path_parity(gpu_ctx, s, cam, tr, 128, 128, "cornell, watertight BVH4", min_frac=0.97)


def _quad_box(accel):
    """A closed box of six inward quads; adjacent quads share their corners
    exactly (corner tuples reused), with coordinates that are not dyadic."""
    lo, hi = (-0.9, -1.6, 0.15), (1.6, 1.3, 1.85)
    s = scenes.Scene(aggregate=capi.AGGREGATE_BVH)
    m = s.add_material(scenes.Material(color=(0.5, 0.5, 0.5)))
    scenes._room_quads(s, lo, hi, m)
    s.accel = accel
    corners = sorted({tuple(np.float32(c) for c in p) for e in s.entities for p in e.points})
    index = {c: k for k, c in enumerate(corners)}
    pos = np.array(corners, np.float32)
    faces = []
    for e in s.entities:
        a, b, c, d = (index[tuple(np.float32(x) for x in p)] for p in e.points)
        faces += [(a, b, c), (a, c, d)]
    return s, pos, np.array(faces), pos[np.array(faces)]


@pytest.mark.parametrize("accel", [capi.ACCEL_BVH4_WATERTIGHT, capi.ACCEL_BVH4])
def test_quad_box_edges(gpu_ctx, accel):
    """Watertight mode on quads: rays from inside aimed exactly at the box's
    corners, shared edges and quad diagonals never escape, and every hit is the
    watertight test of the quad triangle it reports, bit for bit; Atrc's
    Moller-Trumbore quads (the BVH4 mode) let some of the same rays through."""
    s, pos, faces, tris = _quad_box(accel)
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    o, d = _edge_rays(pos, faces, seed=31, n_origins=40)
    n = len(o)
    g = gpu_ctx.trace_rays(o, d, np.zeros(n, np.float32), np.full(n, np.inf, np.float32))
    a = gpu_ctx.trace_rays(o, d, np.zeros(n, np.float32), np.full(n, np.inf, np.float32), any_hit=True)
    escaped, any_escaped = int((g["entity"] < 0).sum()), int((a["entity"] < 0).sum())
    log(f"[watertight] quad box, accel {accel}: {n} corner / edge / diagonal rays from inside, "
        f"escaped {escaped} (any hit {any_escaped})")
    if accel == capi.ACCEL_BVH4:
        assert escaped > 0
        return
    assert escaped == 0 and any_escaped == 0
    # each hit: the watertight test of the reported half (prim 0 = ABC, 1 = ACD) of the reported quad
    tri_index = 2 * g["entity"] + g["prim"]
    for i in range(0, n, 7):
        r = po.triangle_soup_closest(tris[tri_index[i]:tri_index[i] + 1], o[i:i + 1], d[i:i + 1], 0, np.inf,
                                     watertight=True)
        assert r[0][0] == 0
        assert (r[1][0], r[2][0], r[3][0]) == (g["t"][i], g["u"][i], g["v"][i])
    # and the closest hit over all twelve triangles has the same distance
    prim, t, u, v = po.triangle_soup_closest(tris, o, d, 0, np.inf, watertight=True)
    np.testing.assert_array_equal(t, g["t"])
