"""Disney BSDF on the gfx950 path (Atrc material/disney.cpp:18-457) against the
oracle: path for path at 1 spp, and converged against the FP64 oracle."""
import numpy as np
import pytest

from atrc_amd import capi, scenes
from oracle import py_oracle as po

from conftest import log
from test_gpu_render_parity import _path_parity, relmse

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("accel", [capi.ACCEL_ATRC, capi.ACCEL_BVH4])
def test_disney_path_parity(gpu_ctx, accel):
    s, cam, tr = scenes.disney_spheres()
    s.accel = accel
    _path_parity(gpu_ctx, s, cam, tr, 128, 96, f"disney spheres (accel {accel})", min_frac=0.97)


def test_disney_converged(gpu_ctx):
    s, cam, tr = scenes.disney_spheres()
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    W, H = 32, 24
    log("[disney] converged 32x24, GPU 4096 spp vs FP64 oracle 2048 spp")
    g = gpu_ctx.render(scenes.render_params(W, H, 4096, cam, seed=99, **tr))["image"].astype(np.float64)
    r = po.OracleScene(desc, 64).render(scenes.render_params(W, H, 2048, cam, seed=7, **tr), threads=8)["image"]
    e = relmse(g, r)
    log(f"  relMSE {e:.3e}")
    assert e < 3e-3
