"""Batched rendering (agx_render_batch; no counterpart in Atrc): several scenes
in flight at once on worker contexts must give, item for item, the image that
agx_render gives for that scene alone (same per-(pixel, sample) streams, same
film order, so bit-identical), whatever the number of workers the autotuner
picks."""
import numpy as np
import pytest

from atrc_amd import capi, scenes

from conftest import log

pytestmark = pytest.mark.gpu


def _items():
    out = []
    s, cam, tr = scenes.cornell_box()
    out.append((s, scenes.render_params(64, 48, 8, cam, seed=3, **tr)))
    for k, accel in enumerate((capi.ACCEL_BVH4, capi.ACCEL_ATRC, capi.ACCEL_BVH4_WATERTIGHT)):
        s, cam, tr = scenes.procedural_spheres(n_spheres=6 + 4 * k, subdiv=3, seed=10 + k, extent=12.0)
        s.accel = accel
        cam = {"pos": (-8.0, -10.0, 14.0), "dst": (6.0, 6.0, 4.0), "fov": 45.0}
        out.append((s, scenes.render_params(40 + 16 * k, 32 + 8 * k, 4 + 2 * k, cam, seed=20 + k, **tr)))
    s, cam, tr = scenes.volumetric_cornell(grid_n=8)
    out.append((s, scenes.render_params(32, 32, 6, cam, seed=7, **tr)))
    return out


@pytest.mark.parametrize("workers", [3, 0])
def test_batch_matches_single_renders(gpu_ctx, workers):
    items = _items()
    descs, keep, params = [], [], []
    for s, p in items:
        d, k = s.to_desc()
        descs.append(d)
        keep.append(k)
        params.append(p)
    alone = []
    for d, p in zip(descs, params):
        gpu_ctx.upload_scene(d)
        alone.append(gpu_ctx.render(p, aovs=False)["image"])
    log(f"[batch] {len(items)} scenes, workers={workers}")
    imgs, recs, rep = gpu_ctx.render_batch(descs, params, workers=workers)
    log(f"  report {rep}")
    for i, (a, b, r) in enumerate(zip(alone, imgs, recs)):
        assert r["status"] == capi.AGX_OK and r["rays"] > 0
        np.testing.assert_array_equal(a, b, err_msg=f"item {i}")
    assert rep["workers"] >= 1 and rep["rays"] == sum(r["rays"] for r in recs)
    if workers == 0:
        assert rep["tuned"] and rep["tune_rays_per_s"][0] > 0
    else:
        assert rep["workers"] == workers and len({r["worker"] for r in recs}) > 1


def test_batch_rejects_guided_items(gpu_ctx):
    s, cam, tr = scenes.cornell_box()
    d, keep = s.to_desc()
    p = scenes.render_params(16, 16, 1, cam, guide=True, **tr)
    with pytest.raises(capi.AgxError) as e:
        gpu_ctx.render_batch([d], [p], workers=1)
    assert e.value.code == capi.AGX_E_UNSUPPORTED


def test_batch_graph_matches_single_renders(gpu_ctx):
    """agx_render_batch_graph: all items captured into one hipGraph and launched
    once.  Launched twice, with eager renders between the launches: every item of
    every launch is bit-identical to agx_render alone, and its ray count is the
    one the eager batch reports."""
    items = _items()
    s, cam, tr = scenes.cornell_box()
    items.append((s, scenes.render_params(40, 40, 4, cam, seed=5, filter="gaussian", filter_radius=1.5,
                                          filter_alpha=2.0, **tr)))
    descs, keep, params = [], [], []
    for s, p in items:
        d, k = s.to_desc()
        descs.append(d)
        keep.append(k)
        params.append(p)
    alone = []
    for d, p in zip(descs, params):
        gpu_ctx.upload_scene(d)
        alone.append(gpu_ctx.render(p, aovs=False)["image"])
    _, eager_recs, _ = gpu_ctx.render_batch(descs, params, workers=2)
    for launch in range(2):
        imgs, recs, rep = gpu_ctx.render_batch(descs, params, workers=3, graph=True)
        log(f"[batch graph] launch {launch}: {len(items)} scenes, report {rep}")
        for i, (a, b, r, e) in enumerate(zip(alone, imgs, recs, eager_recs)):
            assert r["status"] == capi.AGX_OK
            np.testing.assert_array_equal(a, b, err_msg=f"launch {launch} item {i}")
            assert r["rays"] == e["rays"], (launch, i, r["rays"], e["rays"])
        assert rep["workers"] == 3 and rep["rays"] == sum(r["rays"] for r in recs)
        # eager work between the launches
        gpu_ctx.upload_scene(descs[0])
        np.testing.assert_array_equal(gpu_ctx.render(params[0], aovs=False)["image"], alone[0])


def test_batch_graph_autotuned_and_grouped(gpu_ctx):
    """agx_render_batch_graph with branches = 0: the autotuner times branch
    counts, pass sizes and traversal block sizes on the first group of 8, then
    the batch (here 11 items: two groups, built while the first renders) gives
    every item the image agx_render gives it alone."""
    items = _items()
    s, cam, tr = scenes.cornell_box()
    for k in range(11 - len(items)):
        items.append((s, scenes.render_params(24 + 4 * k, 24, 2 + k % 3, cam, seed=60 + k, **tr)))
    descs, keep, params = [], [], []
    for sc, p in items:
        d, kk = sc.to_desc()
        descs.append(d)
        keep.append(kk)
        params.append(p)
    alone = []
    for d, p in zip(descs, params):
        gpu_ctx.upload_scene(d)
        alone.append(gpu_ctx.render(p, aovs=False)["image"])
    imgs, recs, rep = gpu_ctx.render_batch(descs, params, workers=0, graph=True)
    log(f"[batch graph] tuned: {rep}")
    assert rep["tuned"] and rep["tune_configs"] >= 6
    assert rep["workers"] in (2, 4, 8) and rep["pass_paths_log2"] in (21, 22, 23) and rep["trace_block"] in (128, 256)
    assert rep["groups"] == 2 and rep["build_threads"] >= 1
    for i, (a, b, r) in enumerate(zip(alone, imgs, recs)):
        assert r["status"] == capi.AGX_OK
        np.testing.assert_array_equal(a, b, err_msg=f"item {i}")
    assert rep["rays"] == sum(r["rays"] for r in recs)


@pytest.mark.parametrize("accel", [capi.ACCEL_BVH4, capi.ACCEL_BVH4_WATERTIGHT])
def test_traversal_block_128_matches_256(gpu_ctx, monkeypatch, accel):
    """256-thread traversal blocks (an autotuner choice; AGX_TRACE_BLOCK=256 at
    upload selects them) give the image of the default 128-thread blocks bit
    for bit (both BVH4 modes), with the same ray counts."""
    s, cam, tr = scenes.procedural_spheres(n_spheres=24, subdiv=3, seed=5)
    s.accel = accel
    d, keep = s.to_desc()
    p = scenes.render_params(64, 48, 4, cam, seed=9, **tr)
    gpu_ctx.upload_scene(d)
    a = gpu_ctx.render(p, aovs=False)["image"]
    sa = gpu_ctx.stats()
    monkeypatch.setenv("AGX_TRACE_BLOCK", "256")
    gpu_ctx.upload_scene(d)
    b = gpu_ctx.render(p, aovs=False)["image"]
    sb = gpu_ctx.stats()
    monkeypatch.delenv("AGX_TRACE_BLOCK")
    gpu_ctx.upload_scene(d)
    np.testing.assert_array_equal(a, b)
    assert (sa["closest_rays"], sa["shadow_rays"]) == (sb["closest_rays"], sb["shadow_rays"])
