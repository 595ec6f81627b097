"""Guided path tracing (no counterpart in Atrc): the continuation direction is
drawn from a one-sample mixture of the BSDF and the learned vMF mixture, with
the mixture density in the estimator, so the image must stay the one Atrc's
estimator converges to; on light that enters through a narrow gap the learned
density must cut the error at equal sample count."""
import numpy as np
import pytest

from atrc_amd import scenes
from oracle import py_oracle as po

from conftest import log
from test_gpu_render_parity import relmse

pytestmark = pytest.mark.gpu


def test_guided_cornell_converges_to_the_oracle(gpu_ctx):
    s, cam, tr = scenes.cornell_box()
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    W = H = 32
    log("[guided] cornell 32x32: GPU guided (training on) 4096 spp vs FP64 oracle 4096 spp")
    g = gpu_ctx.render(scenes.render_params(W, H, 4096, cam, seed=99, guide=True, guide_reset=True,
                                            guide_seed=3, **tr))["image"].astype(np.float64)
    st = gpu_ctx.stats()
    log(f"  training samples {st['guide_samples']}, Adam steps {st['guide_steps']}, "
        f"infer {st['guide_infer_ms']:.1f} ms, train {st['guide_train_ms']:.1f} ms")
    assert st["guide_steps"] > 0 and st["guide_samples"] > 0
    r = po.OracleScene(desc, 64).render(scenes.render_params(W, H, 4096, cam, seed=7, **tr), threads=8)["image"]
    e = relmse(g, r)
    log(f"  relMSE {e:.3e}")
    assert np.isfinite(g).all() and e < 3e-3


def relmse_scaled(x, ref):
    """relMSE after scaling both images by 1 / mean(ref): the door-gap room is
    dark, and the 1e-2 floor of relmse would otherwise turn it into plain MSE."""
    k = 1.0 / ref.mean()
    return relmse(x * k, ref * k)


def test_guiding_reduces_error_through_a_door_gap(gpu_ctx):
    """The network trained on a warm-up render (Adam lr 1e-2) and then frozen
    cuts the error at equal sample count.  The gain of one online-trained
    network varies with its initialisation (measured 1.2x to 4.8x over six
    seeds on this 128x128 warm-up), so three networks are trained and the
    median gain must exceed 1.5x."""
    s, cam, tr = scenes.guided_room(n_clutter_tris=50_000, seed=7)
    s.accel = 1
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    W = H = 128
    log(f"[guided] door-gap room {W}x{H}: reference plain 4096 spp")
    ref = gpu_ctx.render(scenes.render_params(W, H, 4096, cam, seed=1, **tr), aovs=False)["image"].astype(np.float64)
    plain = gpu_ctx.render(scenes.render_params(W, H, 256, cam, seed=2, **tr), aovs=False)["image"].astype(np.float64)
    ep = relmse_scaled(plain, ref)
    ratios = []
    for gs in range(3):
        # train on a warm-up render, then render the measured image with the network frozen
        gpu_ctx.render(scenes.render_params(W, H, 256, cam, seed=3, guide=True, guide_reset=True, guide_seed=gs, **tr),
                       aovs=False)
        st = gpu_ctx.stats()
        guided = gpu_ctx.render(scenes.render_params(W, H, 256, cam, seed=4, guide=True, guide_train=False, **tr),
                                aovs=False)["image"].astype(np.float64)
        assert np.isfinite(guided).all()
        eg = relmse_scaled(guided, ref)
        ratios.append(ep / eg)
        log(f"  network seed {gs}: {st['guide_steps']} Adam steps, {st['guide_samples']} samples in the warm-up; "
            f"relMSE at 256 spp: plain {ep:.3e}, guided {eg:.3e} (ratio {ep / eg:.2f})")
    assert float(np.median(ratios)) > 1.5


def test_guided_media_converges_to_the_oracle(gpu_ctx):
    """Guiding in media: scattering points get their own network evaluation
    (k_medium_pass) and sample a one-sample mixture of the learned density and
    the Henyey-Greenstein phase function at every depth; the estimate stays
    Atrc's."""
    s, cam, tr = scenes.volumetric_cornell(grid_n=16)
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    W = H = 32
    log("[guided] volumetric cornell 32x32: GPU guided (training on) 4096 spp vs FP64 oracle 4096 spp")
    g = gpu_ctx.render(scenes.render_params(W, H, 4096, cam, seed=99, guide=True, guide_reset=True,
                                            guide_seed=5, guide_medium_depth=64, **tr))["image"].astype(np.float64)
    st = gpu_ctx.stats()
    log(f"  training samples {st['guide_samples']}, Adam steps {st['guide_steps']}")
    assert st["guide_steps"] > 0 and st["guide_samples"] > 0
    r = po.OracleScene(desc, 64).render(scenes.render_params(W, H, 4096, cam, seed=7, **tr), threads=8)["image"]
    e = relmse(g, r)
    log(f"  relMSE {e:.3e}")
    assert np.isfinite(g).all() and e < 4e-3


def test_guiding_reduces_error_in_fog_behind_a_door_gap(gpu_ctx):
    """The door-gap room filled with thin fog (the volumetric room's fBm
    density scaled to a maximum of 0.05, no slab; 50k clutter): light reaches
    the first room only through the gap, by way of the walls and the fog.
    With the network trained on a warm-up render and then frozen, guided
    rendering cuts the error at equal sample count (measured about 3x)."""
    s, cam, tr = scenes.volumetric_room(n_clutter_tris=50_000, grid_n=32, density_max=0.05, slab=False)
    s.accel = 1
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    W = H = 128
    log(f"[guided] thin-fog room {W}x{H}: reference plain 4096 spp")
    ref = gpu_ctx.render(scenes.render_params(W, H, 4096, cam, seed=1, **tr), aovs=False)["image"].astype(np.float64)
    plain = gpu_ctx.render(scenes.render_params(W, H, 128, cam, seed=2, **tr), aovs=False)["image"].astype(np.float64)
    gpu_ctx.render(scenes.render_params(W, H, 256, cam, seed=3, guide=True, guide_reset=True, **tr), aovs=False)
    guided = gpu_ctx.render(scenes.render_params(W, H, 128, cam, seed=4, guide=True, guide_train=False, **tr),
                            aovs=False)["image"].astype(np.float64)
    ep, eg = relmse_scaled(plain, ref), relmse_scaled(guided, ref)
    log(f"  relMSE at 128 spp: plain {ep:.3e}, guided {eg:.3e} (ratio {ep / eg:.2f})")
    assert np.isfinite(guided).all()
    assert eg * 1.5 < ep


def test_guided_media_keep_the_mean(gpu_ctx):
    """Guiding at medium points at every depth leaves the image mean where
    plain rendering puts it (thin and thicker fog, 16x16-pixel block means of
    8192-spp renders against two plain renders)."""
    W = H = 128
    for dmax in (0.05, 0.5):
        s, cam, tr = scenes.volumetric_room(n_clutter_tris=50_000, grid_n=32, density_max=dmax, slab=False)
        s.accel = 1
        desc, keep = s.to_desc()
        gpu_ctx.upload_scene(desc)
        pa = gpu_ctx.render(scenes.render_params(W, H, 8192, cam, seed=1, **tr), aovs=False)["image"].astype(np.float64)
        pb = gpu_ctx.render(scenes.render_params(W, H, 8192, cam, seed=2, **tr), aovs=False)["image"].astype(np.float64)
        gpu_ctx.render(scenes.render_params(W, H, 256, cam, seed=3, guide=True, guide_reset=True,
                                            guide_medium_depth=16, **tr), aovs=False)
        g = gpu_ctx.render(scenes.render_params(W, H, 8192, cam, seed=4, guide=True, guide_train=False,
                                                guide_medium_depth=16, **tr), aovs=False)["image"].astype(np.float64)
        st = gpu_ctx.stats()
        plain_mean = 0.5 * (pa.mean() + pb.mean())
        log(f"[guided] fog max density {dmax}: mean guided {g.mean():.5e} vs plain {pa.mean():.5e} / {pb.mean():.5e}; "
            f"network rows {st['guide_rows']}")
        assert st["guide_rows"] > 0
        assert abs(g.mean() / plain_mean - 1) < 0.03  # two plain 8192-spp means differ by about 1 %
        blk = lambda x: x.mean(axis=2).reshape(8, 16, 8, 16).mean(axis=(1, 3))
        dev_pp = np.abs(blk(pa) / blk(pb) - 1).mean()
        dev_gp = np.abs(blk(g) / (0.5 * (blk(pa) + blk(pb))) - 1).mean()
        log(f"  block deviation plain-plain {dev_pp:.4f}, guided-plain {dev_gp:.4f}")
        assert dev_gp < 4 * dev_pp
