"""Participating media on the gfx950 path (Atrc trace_std medium branch,
src/tracer/src/render/path_tracing.cpp:64-119; direct_illum.cpp:44-70,
194-253; core/medium/*.cpp).

* Path-for-path against the FP32 oracle on the volumetric Cornell scene
  (homogeneous + heterogeneous media behind invisible_surface boundaries,
  area light + sky): the GPU draws the same two PCG32 streams per path.
* Closed forms on the GPU alone: the absorbing slab (L exp(-sigma_a d)) and
  the albedo-1 furnace (radiance exactly 1 everywhere).
* Converged: GPU vs FP64 oracle at high spp.
"""
import math

import numpy as np
import pytest

from atrc_amd import capi, scenes
from oracle import py_oracle as po

from conftest import log
from test_gpu_render_parity import _path_parity, relmse
from test_oracle_media import _pixel_cos

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("accel", [capi.ACCEL_ATRC, capi.ACCEL_BVH4])
def test_volumetric_cornell_path_parity(gpu_ctx, accel):
    s, cam, tr = scenes.volumetric_cornell(grid_n=16)
    s.accel = accel
    _path_parity(gpu_ctx, s, cam, tr, 128, 128, f"volumetric cornell (homogeneous + heterogeneous media, "
                 f"accel {accel})", min_frac=0.97)


def test_absorbing_slab_gpu(gpu_ctx):
    sa = np.array([0.3, 0.6, 1.2])
    s, cam, tr = scenes.absorbing_slab(sigma_a=tuple(sa), thickness=1.5)
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    W = H = 32
    log("[media] absorbing slab 32x32 64 spp")
    img = gpu_ctx.render(scenes.render_params(W, H, 64, cam, **tr))["image"].astype(np.float64)
    cos = _pixel_cos(W, H, cam["fov"])
    expect = np.stack([(2 * np.exp(-c * 1.5 / cos)).mean(axis=(1, 3)) for c in sa], axis=-1)
    np.testing.assert_allclose(img, expect, rtol=3e-3)


@pytest.mark.parametrize("het", [False, True])
def test_medium_furnace_gpu(gpu_ctx, het):
    s, cam, tr = scenes.medium_furnace(heterogeneous=het)
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    log(f"[media] furnace ({'heterogeneous' if het else 'homogeneous'}) 64x64 256 spp")
    img = gpu_ctx.render(scenes.render_params(64, 64, 256, cam, seed=5, **tr))["image"].astype(np.float64)
    assert np.isfinite(img).all()
    m = img.mean()
    se = img[..., 0].std() / math.sqrt(img[..., 0].size)
    log(f"  mean {m:.5f} (se {se:.5f})")
    assert abs(m - 1.0) < 4 * se + 1e-3


def test_volumetric_cornell_converged(gpu_ctx):
    s, cam, tr = scenes.volumetric_cornell(grid_n=16)
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    W = H = 24
    log("[media] volumetric cornell converged 24x24, GPU 4096 spp vs FP64 oracle 2048 spp")
    g = gpu_ctx.render(scenes.render_params(W, H, 4096, cam, seed=99, **tr))["image"].astype(np.float64)
    r = po.OracleScene(desc, 64).render(scenes.render_params(W, H, 2048, cam, seed=7, **tr), threads=8)["image"]
    e = relmse(g, r)
    log(f"  relMSE {e:.3e}")
    assert e < 3e-3
