"""Pass-size independence of the GPU path tracer.

A render is cut into wavefront passes of at most the context's pass capacity
(whole sample layers when samples splat).  The film accumulates every pass in
sample order and each path's random stream depends only on its (pixel,
sample), so the image must not depend on the cut: a context limited to a few
thousand (or fewer than one sample layer of) paths per pass (AGX_MAX_PATHS, read at its first render) must give
bit for bit the image, albedo and normal planes of a one-pass render.
"""
import numpy as np
import pytest

from atrc_amd import capi, scenes

from conftest import log

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("filt,cap", [("box", 4096), ("gaussian", 4096), ("box", 1024), ("gaussian", 1024)])
def test_many_passes_equal_one_pass(gpu_ctx, monkeypatch, filt, cap):
    """cap 1024 is below the 48x40 layer (1920 paths): passes then cut sample
    layers, and the film takes partial layers that start mid-image."""
    s, cam, tr = scenes.cornell_box()
    desc, keep = s.to_desc()
    kw = {"filter": "gaussian", "filter_radius": 1.5, "filter_alpha": 2.0} if filt == "gaussian" else {}
    p = scenes.render_params(48, 40, 6, cam, seed=17, **kw, **tr)
    gpu_ctx.upload_scene(desc)
    one = gpu_ctx.render(p)
    monkeypatch.setenv("AGX_MAX_PATHS", str(cap))
    with capi.Context(0) as small:
        small.upload_scene(desc)
        many = small.render(p)
        st = small.stats()
    log(f"[passes] {filt}: {st['paths']} paths in passes of at most {cap} against one pass")
    for k in ("image", "albedo", "normal"):
        assert np.array_equal(one[k].view(np.uint32), many[k].view(np.uint32)), k
