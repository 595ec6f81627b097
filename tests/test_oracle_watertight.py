"""The oracle's watertight ray/triangle test (Woop, Benthin and Wald, JCGT 2(1),
2013; the GPU path's AGX_ACCEL_BVH4_WATERTIGHT mode, not an Atrc function),
pinned independently of its own formula:

* dyadic known answers (t and barycentrics exact), both facings;
* agreement with the Moller-Trumbore restatement of Atrc's test
  (utility/triangle_aux.h:38-69), itself pinned in test_oracle_math.py, on
  random rays whose barycentrics keep clear of the triangle's edges;
* watertightness on a closed mesh: rays from inside aimed exactly at every
  vertex and at points on every edge never escape, and an edge shared by two
  triangles meets the ray in at least one of them (the property the test
  exists for; Moller-Trumbore gives no such guarantee).
"""
import numpy as np

from atrc_amd import scenes
from oracle import py_oracle as po


def test_known_answers():
    tri = np.array([[[0, 0, 1], [1, 0, 1], [0, 1, 1]]], np.float32)
    o = np.array([[0.25, 0.25, 0], [0.25, 0.5, 3], [0, 0, 0], [0.5, 0.5, 0], [0.75, 0.5, 0]], np.float32)
    d = np.array([[0, 0, 1], [0, 0, -1], [0, 0, 1], [0, 0, 1], [0, 0, 1]], np.float32)
    prim, t, u, v = po.triangle_soup_closest(tri, o, d, 0, np.inf, watertight=True)
    # inside from below, inside from above, exactly on vertex A, exactly on edge BC, outside
    np.testing.assert_array_equal(prim, [0, 0, 0, 0, -1])
    np.testing.assert_array_equal(t[:4], [1, 2, 1, 1])
    np.testing.assert_array_equal(u[:4], [0.25, 0.25, 0, 0.5])
    np.testing.assert_array_equal(v[:4], [0.25, 0.5, 0, 0.5])
    # the t interval is inclusive and bounds the hit
    prim, t, _, _ = po.triangle_soup_closest(tri, o[:1], d[:1], 1.0, 1.0, watertight=True)
    assert prim[0] == 0 and t[0] == 1
    prim, _, _, _ = po.triangle_soup_closest(tri, o[:1], d[:1], 0.0, np.nextafter(np.float32(1), np.float32(0)),
                                             watertight=True)
    assert prim[0] == -1


def test_agrees_with_moller_trumbore_away_from_edges():
    rng = np.random.default_rng(5)
    n = 3000
    tris = rng.uniform(-1, 1, (n, 3, 3)).astype(np.float32)
    tris[:, :, 2] += 4
    # aim each ray at a barycentric point of its own triangle, well inside or well outside
    bary = rng.dirichlet((1, 1, 1), n)
    outside = rng.random(n) < 0.3
    bary[outside] = np.array([1.4, -0.2, -0.2])
    target = np.einsum("nk,nkc->nc", bary, tris.astype(np.float64))
    o = rng.uniform(-0.5, 0.5, (n, 3))
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ok = np.ones(n, bool)
    hits_wt, hits_mt = [], []
    for i in range(n):
        r_wt = po.triangle_soup_closest(tris[i:i + 1], o[i:i + 1], d[i:i + 1], 0, np.inf, watertight=True)
        r_mt = po.triangle_soup_closest(tris[i:i + 1], o[i:i + 1], d[i:i + 1], 0, np.inf, watertight=False)
        hits_wt.append(r_wt)
        hits_mt.append(r_mt)
    # keep rays whose barycentrics clear every edge by 1e-3 and that meet the plane at
    # more than a grazing angle (both tests are ill-conditioned there)
    t64 = tris.astype(np.float64)
    nrm = np.cross(t64[:, 1] - t64[:, 0], t64[:, 2] - t64[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    cos = np.abs(np.einsum("nc,nc->n", nrm, d))
    for i in range(n):
        ok[i] = (bary[i].min() > 1e-3 or outside[i]) and cos[i] > 0.05
    n_cmp = 0
    for i in np.nonzero(ok)[0]:
        pw, tw, uw, vw = (x[0] for x in hits_wt[i])
        pm, tm, um, vm = (x[0] for x in hits_mt[i])
        assert (pw == 0) == (not outside[i]), i
        assert pw == pm, i
        if pw == 0:
            n_cmp += 1
            assert abs(tw - tm) <= 1e-4 * abs(tm) + 1e-6, (i, tw, tm)
            assert abs(uw - um) <= 1e-3 and abs(vw - vm) <= 1e-3, (i, uw, um, vw, vm)
    assert n_cmp > 1500


def _closed_mesh(subdiv=2, scale=1.7, centre=(0.3, -0.2, 0.9)):
    pos, faces = scenes.icosphere(subdiv)
    pos = (pos * scale + np.asarray(centre)).astype(np.float32)
    return pos, faces, pos[faces]


def test_closed_mesh_has_no_cracks():
    pos, faces, tris = _closed_mesh()
    c = np.array([0.3, -0.2, 0.9], np.float32)
    origins = [c, c + np.float32([0.11, -0.07, 0.05]), c + np.float32([-0.3, 0.2, -0.1])]
    edges = set()
    for f in faces:
        for a, b in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0])):
            edges.add((min(a, b), max(a, b)))
    edges = np.array(sorted(edges))
    # targets: every vertex, and points on every edge computed in float32
    ts = np.float32([0.5, 0.25, 1.0 / 3.0])
    on_edge = [(pos[edges[:, 0]] * (1 - t) + pos[edges[:, 1]] * t).astype(np.float32) for t in ts]
    targets = np.concatenate([pos] + on_edge)
    for o in origins:
        d = (targets - o).astype(np.float32)
        oo = np.broadcast_to(o, d.shape)
        prim, t, _, _ = po.triangle_soup_closest(tris, oo, d, 0, np.inf, watertight=True)
        assert (prim >= 0).all(), f"{(prim < 0).sum()} of {len(prim)} rays escaped the closed mesh"
        # the hit lies on the surface the ray was aimed at (t = 1 up to rounding)
        np.testing.assert_allclose(t, 1.0, rtol=1e-5)


def test_moller_trumbore_leaks_where_watertight_does_not():
    # the same construction from random interior origins: Atrc's test lets rays slip
    # between triangles at edges and vertices, the watertight test never does
    pos, faces, tris = _closed_mesh()
    edges = set()
    for f in faces:
        for a, b in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0])):
            edges.add((min(a, b), max(a, b)))
    edges = np.array(sorted(edges))
    rng = np.random.default_rng(1)
    misses = {True: 0, False: 0}
    for _ in range(8):
        o = (np.float32([0.3, -0.2, 0.9]) + rng.uniform(-0.5, 0.5, 3)).astype(np.float32)
        ts = rng.random(4).astype(np.float32)
        targets = np.concatenate([pos] + [(pos[edges[:, 0]] * (1 - t) + pos[edges[:, 1]] * t).astype(np.float32)
                                          for t in ts])
        d = (targets - o).astype(np.float32)
        for wt in (True, False):
            prim, _, _, _ = po.triangle_soup_closest(tris, np.broadcast_to(o, d.shape), d, 0, np.inf, watertight=wt)
            misses[wt] += int((prim < 0).sum())
    assert misses[True] == 0
    assert misses[False] > 0


def test_shared_edge_is_hit_by_a_neighbour():
    # two triangles sharing the edge (0,0)-(1,1) in the plane z = 2; rays through points
    # on and within a few ulps of the shared edge must hit one of them
    tris = np.array([[[0, 0, 2], [1, 0, 2], [1, 1, 2]], [[0, 0, 2], [1, 1, 2], [0, 1, 2]]], np.float32)
    rng = np.random.default_rng(3)
    s = rng.random(2000).astype(np.float32)
    jitter = (rng.integers(-3, 4, 2000) * np.float32(2 ** -24)).astype(np.float32)
    x = s
    y = (s + jitter).astype(np.float32)
    o = np.stack([x, y, np.zeros_like(x)], 1).astype(np.float32) + np.float32([0.013, -0.021, 0])
    tgt = np.stack([x, y, np.full_like(x, 2)], 1)
    d = (tgt - o).astype(np.float32)
    keep = (x > 1e-3) & (x < 1 - 1e-3)
    prim, _, _, _ = po.triangle_soup_closest(tris, o[keep], d[keep], 0, np.inf, watertight=True)
    assert (prim >= 0).all()
