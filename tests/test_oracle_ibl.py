"""The ibl environment light in the oracle (envir_light/ibl.cpp,
env_sampler.h, texture2d/*), pinned by:

* direction -> texel mapping (phi = atan2(y, x) in [0, 2 pi) -> u,
  theta = acos(z) -> v) on an image whose texels encode their own index,
  nearest sampling: exact known answers;
* the patch sampler: its pdf integrates to 1 over the sphere, each sample's
  pdf equals pdf(direction), E[L / pdf] over samples = the quadrature of L;
* importance: a single bright texel draws the samples towards it;
* a constant ibl renders the white furnace like a native_sky of the same
  radiance (a Lambert sphere of albedo 1 under radiance 1 is 1 everywhere).
"""
import math

import numpy as np
import pytest

from atrc_amd import capi, scenes
from oracle import py_oracle as po


def _scene(env):
    s = scenes.Scene(aggregate=capi.AGGREGATE_NATIVE)
    m = s.add_material(scenes.Material(color=(1.0, 1.0, 1.0)))
    s.add(scenes.Entity(capi.GEOM_SPHERE, m, radius=1.0))
    s.env = env
    return s


def _oracle(env, precision=64):
    s = _scene(env)
    desc, keep = s.to_desc()
    return po.OracleScene(desc, precision), (desc, keep, s)


def index_image(w, h):
    img = np.zeros((h, w, 3), np.float32)
    for y in range(h):
        for x in range(w):
            img[y, x] = (x, y, 1.0)
    return img


def test_direction_to_texel_mapping():
    w, h = 16, 8
    o, _ = _oracle({"image": index_image(w, h), "linear": False})
    dirs, expect = [], []
    rng = np.random.default_rng(1)
    for _ in range(400):
        theta = rng.uniform(0.01, math.pi - 0.01)
        phi = rng.uniform(0.0, 2 * math.pi)
        d = (math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta))
        u, v = phi / (2 * math.pi), theta / math.pi
        x, y = min(w - 1, int(u * w)), min(h - 1, int(v * h))
        # stay clear of texel borders, where float rounding may pick the neighbour
        if abs(u * w - round(u * w)) < 1e-3 or abs(v * h - round(v * h)) < 1e-3:
            continue
        dirs.append(d)
        expect.append((x, y, 1.0))
    got = o.env("radiance", np.array(dirs))
    np.testing.assert_array_equal(got, np.array(expect))
    # +z is the first row, -z the last
    np.testing.assert_array_equal(o.env("radiance", [[0, 0, 1]])[0][1], 0)
    np.testing.assert_array_equal(o.env("radiance", [[0, 0, -1]])[0][1], h - 1)


def _bright_env(w=32, h=16):
    rng = np.random.default_rng(4)
    img = rng.uniform(0.05, 0.2, (h, w, 3)).astype(np.float32)
    img[5, 20] = (400.0, 300.0, 200.0)
    return img


def test_sampler_pdf_is_a_density_and_consistent():
    w, h = 32, 16
    img = _bright_env(w, h)
    o, _ = _oracle({"image": img, "linear": False})
    rng = np.random.default_rng(7)
    # E_uniform[4 pi pdf] = 1
    d = rng.normal(size=(200000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    assert abs((4 * math.pi * o.env("pdf", d)).mean() - 1) < 0.02
    # each sample's pdf is pdf(direction)
    s = o.env("sample", rng.random((100000, 3)))
    np.testing.assert_allclose(np.linalg.norm(s[:, :3], axis=1), 1, rtol=1e-9)
    np.testing.assert_allclose(o.env("pdf", s[:, :3]), s[:, 3], rtol=1e-6)
    # E[lum L / pdf] = integral of lum L, exact for nearest texels: sum of texel x solid angle
    lum = lambda c: c @ np.array([0.2126, 0.7152, 0.0722])
    est = (lum(o.env("radiance", s[:, :3])) / s[:, 3]).mean()
    ys = np.arange(h)
    band = 2 * math.pi / w * (np.cos(math.pi * ys / h) - np.cos(math.pi * (ys + 1) / h))
    exact = float((lum(img.astype(np.float64)) * band[:, None]).sum())
    assert abs(est / exact - 1) < 0.03


def test_importance_sampling_finds_the_bright_texel():
    w, h = 32, 16
    o, _ = _oracle({"image": _bright_env(w, h), "linear": False})
    s = o.env("sample", np.random.default_rng(3).random((4000, 3)))
    d = s[:, :3]
    phi = np.mod(np.arctan2(d[:, 1], d[:, 0]), 2 * math.pi)
    theta = np.arccos(np.clip(d[:, 2], -1, 1))
    x, y = np.minimum(w - 1, (phi / (2 * math.pi) * w).astype(int)), np.minimum(h - 1, (theta / math.pi * h).astype(int))
    near = (np.abs(x - 20) <= 1) & (np.abs(y - 5) <= 1)
    assert near.mean() > 0.5
    # no importance sampling: uniform sphere (pdf 1 / 4 pi everywhere)
    o2, _ = _oracle({"image": _bright_env(w, h), "no_importance_sampling": True})
    np.testing.assert_allclose(o2.env("pdf", d[:100]), 1 / (4 * math.pi), rtol=1e-9)


def test_constant_ibl_furnace_matches_native_sky():
    cam = {"pos": (0.0, -4.0, 0.0), "dst": (0.0, 0.0, 0.0), "up": (0, 0, 1), "fov": 40.0}
    p = scenes.render_params(12, 12, 64, cam, seed=5, min_depth=5, max_depth=40, cont_prob=0.9)
    for env in ({"texel": (1.0, 1.0, 1.0)}, {"checker": (4, (1.0, 1.0, 1.0), (1.0, 1.0, 1.0))}):
        o, _ = _oracle(env)
        img = o.render(p, threads=8)["image"]
        assert abs(img.mean() - 1) < 0.02, env
