"""Wide film filters on the gfx950 path (k_film_splat: each pixel gathers, in a
fixed order, the samples of its neighbourhood that Atrc's apply_image_filter
would splat onto it from the pixel's own render tile) against the oracle,
which splats as Atrc's tiles do (tests/test_oracle_filter.py pins it):

* path for path at 1 spp on the Cornell scene, Gaussian and wide box;
* the half-plane emitter's closed form (the filter's integral across the edge).
"""
import numpy as np
import pytest

from atrc_amd import scenes
from oracle import py_oracle as po

from conftest import log
from test_gpu_render_parity import relmse
from test_oracle_filter import filter_profile

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,r,alpha,tile", [("gaussian", 1.5, 2.0, 32), ("box", 1.2, 0.0, 32),
                                               ("gaussian", 2.3, 0.7, 8)])
def test_filter_path_parity(gpu_ctx, kind, r, alpha, tile):
    s, cam, tr = scenes.cornell_box()
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    p = scenes.render_params(64, 48, 2, cam, seed=77, filter=kind, filter_radius=r, filter_alpha=alpha,
                             task_grid_size=tile, **tr)
    log(f"[filter] cornell 64x48 2 spp, {kind} r={r}, tiles of {tile}")
    g = gpu_ctx.render(p)
    o = po.OracleScene(desc, 32).render(p, threads=8)
    gi, oi = g["image"].astype(np.float64), o["image"]
    rel = np.abs(gi - oi).max(axis=2) / (np.abs(oi).max(axis=2) + 1e-2)
    frac = float((rel <= 1e-3).mean())
    log(f"  per-pixel agreement {frac:.5f}; relMSE {relmse(gi, oi):.3e}")
    assert np.isfinite(gi).all() and frac >= 0.97
    np.testing.assert_allclose(g["albedo"], o["albedo"], atol=1e-3)


@pytest.mark.parametrize("kind,r,alpha,W", [("gaussian", 1.5, 2.0, 16), ("box", 2.3, 0.0, 4)])
def test_filter_edge_closed_form(gpu_ctx, kind, r, alpha, W):
    s, cam, tr = scenes.half_plane_emitter()
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    rr = r + (3e-4 if kind == "box" else 0.0)
    p = scenes.render_params(W, 6, 2048, cam, seed=5, filter=kind, filter_radius=rr, filter_alpha=alpha, **tr)
    img = gpu_ctx.render(p, aovs=False)["image"][..., 0].astype(np.float64)
    expect = filter_profile(kind, r, alpha, W, lit_right=img[:, -1].mean() > img[:, 0].mean())
    log(f"[filter] half-plane {kind} r={r} W={W}: max |row - closed form| "
        f"{np.abs(img.mean(axis=0) - expect).max():.4f}")
    np.testing.assert_allclose(img.mean(axis=0), expect, atol=0.01)
