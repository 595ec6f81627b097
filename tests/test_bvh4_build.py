"""Structure of the production BVH4 (AGX_ACCEL_BVH4) built at upload, checked
on the CPU through agx_debug_bvh4: a tree over every primitive exactly once,
leaves of at most 4 primitives, every child box -- decoded from its bytes as
the traversal decodes it, in FP32 -- enclosing its subtree, and
primitive records that carry the mesh triangles of Atrc's own triangle BVH
(src/tracer/src/core/geometry/triangle_bvh.cpp:24-27, 225-304) unchanged.
"""
import numpy as np
import pytest

from atrc_amd import capi, scenes

EMPTY = 0xFFFFFFFF
LEAF = 0x80000000
ANALYTIC = 0xFFFFFFFF


def _decode(nodes):
    """Child boxes of 64-byte nodes as b4_children decodes them: origin +
    float(byte) * 2^e, an FP32 multiply then an FP32 add."""
    u = nodes.view(np.uint32)
    eb = u[:, 0, 3]
    org = nodes[:, 0, :3]
    words = [u[:, 2, 0], u[:, 2, 1], u[:, 2, 2], u[:, 2, 3], u[:, 3, 0], u[:, 3, 1]]
    lo = np.zeros((len(nodes), 4, 3), np.float32)
    hi = np.zeros((len(nodes), 4, 3), np.float32)
    for a in range(3):
        step = (((eb >> np.uint32(8 * a)) & np.uint32(0xFF)) << np.uint32(23)).astype(np.uint32).view(np.float32)
        for k in range(4):
            ql = ((words[2 * a] >> np.uint32(8 * k)) & np.uint32(0xFF)).astype(np.float32)
            qh = ((words[2 * a + 1] >> np.uint32(8 * k)) & np.uint32(0xFF)).astype(np.float32)
            lo[:, k, a] = org[:, a] + ql * step
            hi[:, k, a] = org[:, a] + qh * step
    return u[:, 1, :], lo, hi


def test_decode_is_fp32_multiply_then_add():
    # a hand-made node: origin (1, -2, 0.5), steps 2^-3, 2^0, 2^-8; child 2 bytes chosen
    n = np.zeros((1, 4, 4), np.float32)
    u = n.view(np.uint32)
    n[0, 0, :3] = [1.0, -2.0, 0.5]
    u[0, 0, 3] = (124) | (127 << 8) | (119 << 16)
    u[0, 1, :] = [0xFFFFFFFF, 0xFFFFFFFF, 5, 0xFFFFFFFF]
    u[0, 2, :] = [8 << 16, 255 << 16, 3 << 16, 4 << 16]
    u[0, 3, :2] = [0, 128 << 16]
    codes, lo, hi = _decode(n)
    assert codes[0, 2] == 5
    np.testing.assert_array_equal(lo[0, 2], [2.0, 1.0, 0.5])
    np.testing.assert_array_equal(hi[0, 2], [1.0 + 255 / 8, 2.0, 1.0])


def _walk(nodes, prims):
    codes, dlo, dhi = _decode(nodes)
    seen_nodes = np.zeros(len(nodes), np.int32)
    seen_prims = np.zeros(len(prims), np.int32)
    leaf_sizes = []
    # subtree bounds of every node, bottom-up via explicit DFS order
    order, stack = [], [0]
    while stack:
        n = stack.pop()
        order.append(n)
        for k in range(4):
            c = int(codes[n, k])
            if c != EMPTY and not (c & LEAF):
                seen_nodes[c] += 1
                stack.append(c)
    A = prims[:, 0, :3].astype(np.float64)
    B = A + prims[:, 1, :3]
    Cc = A + prims[:, 2, :3]
    slot = prims[:, 1, 3].view(np.uint32)
    tri = slot != ANALYTIC
    lo_p = np.minimum(np.minimum(A, B), Cc)
    hi_p = np.maximum(np.maximum(A, B), Cc)
    sub_lo = np.full((len(nodes), 3), np.inf)
    sub_hi = np.full((len(nodes), 3), -np.inf)
    for n in reversed(order):
        for k in range(4):
            c = int(codes[n, k])
            if c == EMPTY:
                continue
            box_lo = dlo[n, k].astype(np.float64)
            box_hi = dhi[n, k].astype(np.float64)
            if c & LEAF:
                first, cnt = c & 0x07FFFFFF, ((c >> 27) & 15) + 1
                leaf_sizes.append(cnt)
                idx = np.arange(first, first + cnt)
                seen_prims[idx] += 1
                t = idx[tri[idx]]
                if len(t):
                    assert (lo_p[t] >= box_lo).all() and (hi_p[t] <= box_hi).all()
                    sub_lo[n] = np.minimum(sub_lo[n], lo_p[t].min(axis=0))
                    sub_hi[n] = np.maximum(sub_hi[n], hi_p[t].max(axis=0))
            else:
                assert (sub_lo[c] >= box_lo).all() and (sub_hi[c] <= box_hi).all()
                sub_lo[n] = np.minimum(sub_lo[n], sub_lo[c])
                sub_hi[n] = np.maximum(sub_hi[n], sub_hi[c])
    assert seen_nodes[0] == 0 and (seen_nodes[1:] == 1).all()
    assert (seen_prims == 1).all()
    assert max(leaf_sizes) <= 4
    return slot


@pytest.mark.parametrize("which", ["spheres", "cornell"])
def test_bvh4_structure(which):
    if which == "spheres":
        s, _, _ = scenes.procedural_spheres(n_spheres=24, subdiv=3, seed=3, extent=20.0)
    else:
        s, _, _ = scenes.cornell_box()
    s.accel = capi.ACCEL_BVH4
    desc, keep = s.to_desc()
    nodes, prims = capi.debug_bvh4(desc)
    slot = _walk(nodes, prims)
    ent = prims[:, 0, 3].view(np.int32)
    # one record per analytic entity and per mesh triangle
    mesh_ents = [i for i, e in enumerate(s.entities) if e.tris is not None]
    n_tris = sum(len(s.entities[i].tris) // 3 for i in mesh_ents)
    assert len(prims) == n_tris + len(s.entities) - len(mesh_ents)
    analytic = slot == ANALYTIC
    assert sorted(ent[analytic].tolist()) == [i for i in range(len(s.entities)) if i not in mesh_ents]
    # mesh triangles: Atrc's triangle records (a, b - a, c - a), slots contiguous per
    # entity in entity order; a mesh that emits nothing keeps its input triangle
    # order (the BVH4 needs no Atrc tree for it)
    base = 0
    for i in mesh_ents:
        tris = s.entities[i].tris
        n = len(tris) // 3
        ref_nodes, ref_prims, _ = capi.debug_triangle_bvh(tris)
        ref_prims = ref_prims.reshape(n, 3, 4)
        ids = ref_prims[:, 0, 3].view(np.uint32)
        if not any(s.entities[i].emit):
            ref_prims = ref_prims[np.argsort(ids)]
        sel = np.where(~analytic & (ent == i))[0]
        assert len(sel) == n
        local = slot[sel].astype(np.int64) - base
        assert sorted(local.tolist()) == list(range(n))
        np.testing.assert_array_equal(prims[sel, :, :3], ref_prims[local][:, :, :3])
        base += n
    summ = capi.debug_scene_summary(desc)
    assert summ["bvh4_nodes"] == len(nodes) and summ["bvh4_depth"] >= 1
    assert summ["bvh4_depth"] * 3 + 1 <= 192


def test_bvh4_parallel_build_is_reproducible():
    """Above 65,536 references the builder hands subtrees to worker threads,
    each writing its own slot range of the binary-node array.  The tree may
    not depend on which thread finishes first: 81,920 triangles (16
    subdivision-4 spheres) built three times give the same nodes and records
    byte for byte, and the tree is well formed."""
    s, _, _ = scenes.procedural_spheres(n_spheres=16, subdiv=4, seed=9, extent=30.0)
    s.accel = capi.ACCEL_BVH4
    desc, keep = s.to_desc()
    nodes, prims = capi.debug_bvh4(desc)
    assert len(prims) >= 65536 + 1
    for _ in range(2):
        n2, p2 = capi.debug_bvh4(desc)
        assert np.array_equal(nodes.view(np.uint32), n2.view(np.uint32))
        assert np.array_equal(prims.view(np.uint32), p2.view(np.uint32))
    slot = _walk(nodes, prims)
    assert len(slot) == len(prims)


def test_bvh4_is_shallower_than_atrc_bvh():
    s, _, _ = scenes.procedural_spheres(n_spheres=24, subdiv=3, seed=3, extent=20.0)
    s.aggregate = capi.AGGREGATE_BVH
    s.accel = capi.ACCEL_ATRC
    desc, keep = s.to_desc()
    atrc = capi.debug_scene_summary(desc)
    s.accel = capi.ACCEL_BVH4
    desc, keep = s.to_desc()
    summ = capi.debug_scene_summary(desc)
    assert summ["bvh4_depth"] < atrc["max_blas_depth"] + atrc["max_tlas_depth"]
    # the BVH4 mode builds no Atrc triangle BVH for meshes that emit nothing
    assert summ["blas_nodes"] == 0 and summ["blas_prims"] == atrc["blas_prims"]


def test_unknown_accel_rejected():
    s, _, _ = scenes.cornell_box()
    s.accel = 7
    desc, keep = s.to_desc()
    with pytest.raises(capi.AgxError):
        capi.debug_scene_summary(desc)


def test_watertight_records_hold_the_mesh_vertices():
    """AGX_ACCEL_BVH4_WATERTIGHT: the same tree, mesh triangles as (A, B, C) with
    the mesh's own vertex coordinates (shared edges stay bit-identical)."""
    s, _, _ = scenes.procedural_spheres(n_spheres=6, subdiv=2, seed=4, extent=10.0, ground=False)
    s.accel = capi.ACCEL_BVH4
    n4, p4 = capi.debug_bvh4(s.to_desc()[0])
    s.accel = capi.ACCEL_BVH4_WATERTIGHT
    desc, keep = s.to_desc()
    nw, pw = capi.debug_bvh4(desc)
    np.testing.assert_array_equal(n4, nw)
    slot4, slotw = p4[:, 1, 3].view(np.uint32), pw[:, 1, 3].view(np.uint32)
    np.testing.assert_array_equal(slot4, slotw)
    np.testing.assert_array_equal(p4[:, 0, 3].view(np.int32), pw[:, 0, 3].view(np.int32))
    mesh_ents = [i for i, e in enumerate(s.entities) if e.tris is not None]
    ent = pw[:, 0, 3].view(np.int32)
    for i in mesh_ents:
        tris = s.entities[i].tris
        assert not any(s.entities[i].emit)   # slots then follow the input triangle order
        sel = np.where((slotw != ANALYTIC) & (ent == i))[0]
        base = int(slotw[sel].min())
        ids = slotw[sel].astype(np.int64) - base
        verts = np.asarray(tris, np.float32)[:, :3].reshape(-1, 3, 3)
        np.testing.assert_array_equal(pw[sel, :, :3], verts[ids])


def test_watertight_quads_are_two_triangles_over_their_corners():
    """AGX_ACCEL_BVH4_WATERTIGHT: every quad is two primitives, (A, B, C) and
    (A, C, D) over the descriptor's own corners, so quads sharing an edge
    share bit-identical endpoints; spheres stay single analytic primitives."""
    s = scenes.Scene(aggregate=capi.AGGREGATE_NATIVE)
    m = s.add_material(scenes.Material())
    scenes._room_quads(s, (0.1, 0.2, 0.3), (1.7, 2.9, 1.3), m)
    s.add(scenes.Entity(capi.GEOM_SPHERE, m, radius=0.3, local_to_world=(1, 0, 0, 0.9, 0, 1, 0, 1.5, 0, 0, 1, 0.8)))
    s.accel = capi.ACCEL_BVH4_WATERTIGHT
    desc, keep = s.to_desc()
    nodes, prims = capi.debug_bvh4(desc)
    slot = prims[:, 1, 3].view(np.uint32)
    ent = prims[:, 0, 3].view(np.int32)
    assert (slot == ANALYTIC).sum() == 1 and ent[slot == ANALYTIC][0] == 6
    for i, e in enumerate(s.entities[:6]):
        P = np.asarray(e.points, np.float32)
        for code, (b, c) in ((0xFFFFFFFE, (1, 2)), (0xFFFFFFFD, (2, 3))):
            k = np.nonzero((ent == i) & (slot == code))[0]
            assert len(k) == 1
            np.testing.assert_array_equal(prims[k[0], :, :3], P[[0, b, c]])
