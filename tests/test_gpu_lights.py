"""Area lights on the GPU against closed forms and the oracle.

The scenes are the closed-form ones of test_oracle_lights.py (surface
furnaces of quads and of a mesh; sphere and disk lights inside a Lambert
shell), rendered by the gfx950 path tracer:

* converged: the pixel mean equals the closed form within 4 standard errors
  (+3e-4 for the shell's facets);
* path for path at 1 spp against the FP32 oracle (same random stream per
  pixel and sample), image and every AOV plane.

Area-light NEE and MIS at surface vertices for quad, mesh, sphere and disk
emitters: quad.cpp:80-121, sphere.cpp:81-110, disk.cpp,
geometry_to_diffuse_light.cpp:14-90, direct_illum.cpp:11-192.
"""
import numpy as np
import pytest

from atrc_amd import capi, scenes
from oracle import py_oracle as po

from conftest import log

pytestmark = pytest.mark.gpu

CASES = [("furnace", "quads"), ("furnace", "mesh"), ("shell", "sphere"), ("shell", "disk")]


def _scene(which, kind):
    return scenes.emitting_enclosure(kind) if which == "furnace" else scenes.light_in_shell(kind)


@pytest.mark.parametrize("which,kind", CASES)
@pytest.mark.parametrize("accel", [capi.ACCEL_ATRC, capi.ACCEL_BVH4])
def test_closed_form(gpu_ctx, which, kind, accel):
    s, cam, tr, expected = _scene(which, kind)
    s.accel = accel
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    log(f"[lights] {which}/{kind} accel {accel}: 128x128 512 spp")
    img = gpu_ctx.render(scenes.render_params(128, 128, 512, cam, seed=31, **tr))["image"].astype(np.float64)
    m = float(img.mean())
    se = float(img.std()) / np.sqrt(img.size)
    log(f"  mean {m:.6f} expected {expected:.6f} ({(m / expected - 1):+.2e}, stderr {se / expected:.1e})")
    assert abs(m - expected) <= 4 * se + 3e-4 * expected


@pytest.mark.parametrize("which,kind", CASES)
def test_path_parity(gpu_ctx, which, kind):
    s, cam, tr, expected = _scene(which, kind)
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    p = scenes.render_params(64, 64, 1, cam, seed=32, **tr)
    log(f"[lights] {which}/{kind}: path parity 64x64 1 spp")
    g = gpu_ctx.render(p)
    r = po.OracleScene(desc, 32).render(p, threads=8)
    gi, ri = g["image"].astype(np.float64), r["image"]
    rel = np.abs(gi - ri).max(axis=2) / (np.abs(ri).max(axis=2) + 1e-2)
    frac = float((rel <= 1e-3).mean())
    log(f"  per-path agreement {frac:.5f}")
    assert frac >= 0.98
    np.testing.assert_allclose(g["albedo"], r["albedo"], atol=1e-4)
    np.testing.assert_allclose(g["normal"], r["normal"], atol=1e-4)
    np.testing.assert_allclose(g["denoise"], r["denoise"], atol=1e-6)


@pytest.mark.parametrize("kind", ["quads", "mesh"])
def test_furnace_without_mis(gpu_ctx, kind):
    """trace_nomis on the surface furnaces: L (1 - rho^D) / (1 - rho) at max depth D."""
    D = 4
    s, cam, tr, _ = scenes.emitting_enclosure(kind, max_depth=D)
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    expected = (1.0 - 0.5 ** D) / 0.5
    img = gpu_ctx.render(scenes.render_params(128, 128, 512, cam, seed=33, use_mis=False, **tr))["image"]
    img = img.astype(np.float64)
    m = float(img.mean())
    se = float(img.std()) / np.sqrt(img.size)
    log(f"[lights] trace_nomis furnace/{kind}: mean {m:.6f} expected {expected:.6f}")
    assert abs(m - expected) <= 4 * se + 3e-4 * expected


def test_double_sided_closed_forms(gpu_ctx):
    """double_sided on the GPU: a disk light facing away from the receivers
    lights them from its back side (the shell closed form), and outward-facing
    double-sided furnace walls read the surface-furnace closed form."""
    cases = [scenes.light_in_shell("disk", double_sided=True, face_up=True),
             scenes.emitting_enclosure("quads", outward_double_sided=True),
             scenes.emitting_enclosure("mesh", outward_double_sided=True)]
    for k, (s, cam, tr, expected) in enumerate(cases):
        s.accel = capi.ACCEL_BVH4
        desc, keep = s.to_desc()
        gpu_ctx.upload_scene(desc)
        img = gpu_ctx.render(scenes.render_params(128, 128, 512, cam, seed=34 + k, **tr))["image"].astype(np.float64)
        m = float(img.mean())
        se = float(img.std()) / np.sqrt(img.size)
        log(f"[lights] double-sided case {k}: mean {m:.6f} expected {expected:.6f}")
        assert abs(m - expected) <= 4 * se + 3e-4 * expected


@pytest.mark.parametrize("accel", [capi.ACCEL_ATRC, capi.ACCEL_BVH4])
def test_double_sided_path_parity(gpu_ctx, accel):
    """The delta/lens scene with every entity double-sided (walls, boxes, glass,
    mirror and the quad, sphere and disk lights), path for path against the oracle."""
    s, cam, tr = scenes.cornell_delta(True)
    for e in s.entities:
        e.double_sided = True
    s.accel = accel
    desc, keep = s.to_desc()
    gpu_ctx.upload_scene(desc)
    p = scenes.render_params(96, 96, 1, cam, seed=35, **tr)
    log(f"[lights] double-sided delta scene, accel {accel}")
    g = gpu_ctx.render(p)
    r = po.OracleScene(desc, 32).render(p, threads=8)
    gi, ri = g["image"].astype(np.float64), r["image"]
    rel = np.abs(gi - ri).max(axis=2) / (np.abs(ri).max(axis=2) + 1e-2)
    frac = float((rel <= 1e-3).mean())
    log(f"  per-path agreement {frac:.5f}")
    assert frac >= 0.98
    np.testing.assert_allclose(g["normal"], r["normal"], atol=1e-4)
