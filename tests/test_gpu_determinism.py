"""Run-to-run determinism of the GPU path tracer.

The bounce queues are built without atomics, in path order, and the guiding
network's training samples reach the ring in record order; every floating-point
reduction runs in a fixed order.  Two renders with the same parameters
therefore give the same image bit for bit: a plain render (image and AOV
planes), and a guided render that trains a network from the same reset state
and then renders with it.
"""
import numpy as np
import pytest

from atrc_amd import scenes

from conftest import log

pytestmark = pytest.mark.gpu


def test_plain_render_is_bitwise_reproducible(gpu_ctx):
    s, cam, tr = scenes.procedural_spheres(n_spheres=12, subdiv=3, seed=5, extent=12.0)
    s.accel = 1
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    p = scenes.render_params(96, 64, 24, cam, seed=11, **tr)
    a = gpu_ctx.render(p, aovs=True)
    b = gpu_ctx.render(p, aovs=True)
    for k in ("image", "albedo", "normal"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    st = gpu_ctx.stats()
    log(f"[determinism] plain 96x64x24: {st['closest_rays']} closest + {st['shadow_rays']} shadow rays, "
        "two renders bit-identical")


def test_guided_training_is_bitwise_reproducible(gpu_ctx):
    s, cam, tr = scenes.guided_room(n_clutter_tris=20_000, seed=3)
    s.accel = 1
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    imgs = []
    for _ in range(2):
        gpu_ctx.render(scenes.render_params(64, 64, 32, cam, seed=5, guide=True, guide_reset=True, guide_seed=9,
                                            **tr), aovs=False)
        imgs.append(gpu_ctx.render(scenes.render_params(64, 64, 16, cam, seed=6, guide=True, guide_train=False,
                                                        **tr), aovs=False)["image"])
    st = gpu_ctx.stats()
    assert st["guide_steps"] > 0
    assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
    log(f"[determinism] guided 64x64: trained twice from the same reset ({st['guide_steps']} Adam steps), "
        "renders bit-identical")
