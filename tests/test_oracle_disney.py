"""Pins the oracle's Disney BSDF (Atrc material/disney.cpp:18-457) against
textbook forms of each lobe, restated here independently:

* diffuse: Burley 2012, f = C/pi (1 + (F_D90 - 1)(1 - cos_i)^5)(1 + (F_D90 - 1)(1 - cos_o)^5),
  F_D90 = 0.5 + 2 roughness cos_d^2 (Atrc writes it as Lambert + retro-reflection);
* metallic specular: GGX / Trowbridge-Reitz D, separable Smith G1 = 2 / (1 + sqrt(1 + a^2 tan^2)),
  Schlick Fresnel with R0 = base colour;
* clearcoat: Burley GTR1 D (normalised: integral of D cos(theta_h) = 1), Schlick R0 = 0.04,
  Smith G with alpha 0.25;
* transmission: Walter et al. 2007 rough-dielectric BTDF in pbrt-v3's radiance form;
* sampling: a sample returns eval / pdf of its direction; where Atrc's sampler
  and pdf agree (rays inside the surface, isotropic metal) sampled directions
  follow pdf (chi-square).

Quirks reproduced and documented, not pinned against Atrc output (there is
none): the diffuse pick is overwritten by the specular / transmission chain
(disney.cpp:567-587) and pdf_specular uses |sin(phi_o)| (disney.cpp:335-336).
"""
import math

import numpy as np
import pytest

from atrc_amd import capi
from oracle import py_oracle as po


def disney(**kw):
    m = capi.agx_material_desc()
    m.kind = capi.MAT_DISNEY
    m.color[:] = kw.get("base_color", (0.8, 0.5, 0.3))
    m.color2[:] = kw.get("specular_scale", (1.0, 1.0, 1.0))
    m.roughness = kw.get("roughness", 0.5)
    m.anisotropic = kw.get("anisotropic", 0.0)
    m.ior = kw.get("ior", 1.5)
    m.metallic = kw.get("metallic", 0.0)
    m.transmission = kw.get("transmission", 0.0)
    m.transmission_roughness = kw.get("transmission_roughness", m.roughness)
    m.specular_tint = kw.get("specular_tint", 0.0)
    m.sheen = kw.get("sheen", 0.0)
    m.sheen_tint = kw.get("sheen_tint", 0.0)
    m.clearcoat = kw.get("clearcoat", 0.0)
    m.clearcoat_gloss = kw.get("clearcoat_gloss", 1.0)
    return m


def f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


def rand_dirs(rng, n, upper=True):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v[:, 2] = np.abs(v[:, 2]) if upper else -np.abs(v[:, 2])
    return v[np.abs(v[:, 2]) > 0.05]


def ev(m, wi, wo):
    return po.bsdf(m, 0, wi=wi, wo=wo)[:3]


def pdf(m, wi, wo):
    return po.bsdf(m, 1, wi=wi, wo=wo)[0]


def smith_g1(a, tan):
    return 2 / (1 + math.sqrt(1 + a * a * tan * tan))


def ggx_d(a, cos_h):
    return a * a / (math.pi * ((a * a - 1) * cos_h * cos_h + 1) ** 2)


def fresnel_dielectric(cos_i, eta_out, eta_in):
    """Unpolarised Fresnel reflectance for light arriving from the eta_out side."""
    if cos_i < 0:
        eta_out, eta_in, cos_i = eta_in, eta_out, -cos_i
    sin_t = eta_out / eta_in * math.sqrt(max(0.0, 1 - cos_i * cos_i))
    if sin_t >= 1:
        return 1.0
    cos_t = math.sqrt(1 - sin_t * sin_t)
    rs = (eta_out * cos_i - eta_in * cos_t) / (eta_out * cos_i + eta_in * cos_t)
    rp = (eta_in * cos_i - eta_out * cos_t) / (eta_in * cos_i + eta_out * cos_t)
    return 0.5 * (rs * rs + rp * rp)


def test_burley_diffuse():
    rng = np.random.default_rng(1)
    r = 0.6
    C = f32((0.8, 0.5, 0.3))
    m = disney(base_color=tuple(C), specular_scale=(0, 0, 0), roughness=r)
    for wi, wo in zip(rand_dirs(rng, 40), rand_dirs(rng, 40)):
        h = (wi + wo) / np.linalg.norm(wi + wo)
        cd = float(wi @ h)
        f90 = 0.5 + 2 * f32(r) * cd * cd
        expect = C / math.pi * (1 + (f90 - 1) * (1 - wi[2]) ** 5) * (1 + (f90 - 1) * (1 - wo[2]) ** 5)
        np.testing.assert_allclose(ev(m, wi, wo), expect, rtol=1e-9)
    # below the surface (no transmission): black
    assert not ev(m, np.array([0.3, 0.1, -0.9]), np.array([0, 0, 1.0])).any()


def test_metallic_ggx_with_schlick():
    rng = np.random.default_rng(2)
    r = 0.45
    a = f32(r) ** 2
    C = f32((0.9, 0.6, 0.2))
    m = disney(base_color=tuple(C), metallic=1.0, roughness=r)
    for wi, wo in zip(rand_dirs(rng, 40), rand_dirs(rng, 40)):
        h = (wi + wo) / np.linalg.norm(wi + wo)
        cd = float(wi @ h)
        F = C + (1 - C) * (1 - cd) ** 5
        tan = lambda w: math.sqrt(1 - w[2] ** 2) / w[2]
        G = smith_g1(a, tan(wi)) * smith_g1(a, tan(wo))
        expect = F * ggx_d(a, h[2]) * G / (4 * wi[2] * wo[2])
        np.testing.assert_allclose(ev(m, wi, wo), expect, rtol=1e-8)


def test_clearcoat_gtr1():
    gloss = 0.3
    alpha = max((0.1 * (1 - gloss)) ** 2, 1e-4)
    # GTR1 is normalised: integral over the hemisphere of D(h) cos(theta_h) = 1
    from scipy.integrate import quad

    D = lambda c: (alpha ** 2 - 1) / (math.pi * math.log(alpha ** 2) * (1 + (alpha ** 2 - 1) * c * c))
    assert quad(lambda c: 2 * math.pi * D(c) * c, 0, 1, points=[0.999, 0.9999], limit=400)[0] == \
        pytest.approx(1.0, rel=1e-6)
    m = disney(base_color=(0, 0, 0), specular_scale=(0, 0, 0), clearcoat=1.0, clearcoat_gloss=gloss)
    alpha32 = f32(alpha)
    D32 = lambda c: (alpha32 ** 2 - 1) / (math.pi * math.log(alpha32 ** 2) * (1 + (alpha32 ** 2 - 1) * c * c))
    rng = np.random.default_rng(3)
    for wi, wo in zip(rand_dirs(rng, 30), rand_dirs(rng, 30)):
        h = (wi + wo) / np.linalg.norm(wi + wo)
        cd = float(wi @ h)
        F = 0.04 + 0.96 * (1 - cd) ** 5
        tan = lambda w: math.sqrt(1 - w[2] ** 2) / w[2]
        G = smith_g1(0.25, tan(wi)) * smith_g1(0.25, tan(wo))
        expect = D32(h[2]) * F * G / (4 * wi[2] * wo[2])
        np.testing.assert_allclose(ev(m, wi, wo), [expect] * 3, rtol=1e-7)


def test_rough_transmission_walter():
    rng = np.random.default_rng(4)
    r, ior = 0.5, 1.45
    a = f32(r) ** 2
    C = f32((0.81, 0.64, 0.49))
    m = disney(base_color=tuple(C), transmission=1.0, roughness=r, ior=ior, specular_scale=(0, 0, 0))
    eta_b = f32(ior)
    n_checked = 0
    for wo, wi in zip(rand_dirs(rng, 200, upper=True), rand_dirs(rng, 200, upper=False)):
        eta = eta_b                       # wo outside: eta = eta_inside / eta_outside
        wh = wo + wi * eta
        wh /= np.linalg.norm(wh)
        if wh[2] < 0:
            wh = -wh
        if (wo @ wh > 0) != (wo[2] > 0) or (wi @ wh > 0) != (wi[2] > 0):
            continue
        F = fresnel_dielectric(float(wo @ wh), 1.0, eta_b)
        tan = lambda w: math.sqrt(1 - w[2] ** 2) / w[2]
        G = smith_g1(a, tan(wi)) * smith_g1(a, tan(wo))
        sd = float(wo @ wh + eta * (wi @ wh))
        val = (1 - F) * ggx_d(a, wh[2]) * G * abs(wi @ wh) * abs(wo @ wh) / (abs(wi[2] * wo[2]) * sd * sd)
        np.testing.assert_allclose(ev(m, wi, wo), np.sqrt(C) * val, rtol=1e-7)
        n_checked += 1
    assert n_checked > 20


@pytest.mark.parametrize("params", [
    dict(metallic=0.3, roughness=0.4, sheen=0.5, sheen_tint=0.5, clearcoat=0.7, clearcoat_gloss=0.6,
         specular_tint=0.3, anisotropic=0.5),
    dict(metallic=0.0, transmission=0.8, roughness=0.3, ior=1.6),
])
def test_sample_returns_eval_and_pdf(params):
    m = disney(**params)
    rng = np.random.default_rng(5)
    for wo in np.concatenate([rand_dirs(rng, 20), rand_dirs(rng, 20, upper=False)]):
        for u, v, w in rng.random((10, 3)):
            s = po.bsdf(m, 2, wo=wo, u=u, v=v, w=w)
            d, f, p = s[:3], s[3:6], s[6]
            assert s[7] == 0.0            # never a delta lobe
            if not d.any():
                continue
            np.testing.assert_allclose(f, ev(m, d, wo), rtol=1e-12)
            assert p == pytest.approx(pdf(m, d, wo), rel=1e-12)


def _chi2(dirs, pdf_fn, n_theta=8, n_phi=12):
    """Chi-square of sampled directions against pdf over (cos theta, phi) cells of the sphere."""
    ct = np.clip(dirs[:, 2], -1, 1)
    ph = np.mod(np.arctan2(dirs[:, 1], dirs[:, 0]), 2 * math.pi)
    ce = np.linspace(-1, 1, 2 * n_theta + 1)
    pe = np.linspace(0, 2 * math.pi, n_phi + 1)
    hist, _, _ = np.histogram2d(ct, ph, [ce, pe])
    # expected mass per cell by midpoint quadrature on a finer grid
    k = 6
    exp = np.zeros_like(hist)
    for i in range(len(ce) - 1):
        for j in range(len(pe) - 1):
            cs = ce[i] + (np.arange(k) + 0.5) / k * (ce[i + 1] - ce[i])
            ps = pe[j] + (np.arange(k) + 0.5) / k * (pe[j + 1] - pe[j])
            acc = 0.0
            for c in cs:
                s = math.sqrt(max(0.0, 1 - c * c))
                for p in ps:
                    acc += pdf_fn(np.array([s * math.cos(p), s * math.sin(p), c]))
            exp[i, j] = acc / (k * k) * (ce[i + 1] - ce[i]) * (pe[j + 1] - pe[j])
    exp *= len(dirs) / max(exp.sum(), 1e-30)
    sel = exp > 5
    return float(np.sum((hist[sel] - exp[sel]) ** 2 / exp[sel])), int(sel.sum())


def test_sampling_matches_pdf_inside_and_for_metal():
    rng = np.random.default_rng(6)
    # (a) rays inside a rough dielectric: transmission + internal reflection, exact mixture
    m = disney(transmission=1.0, roughness=0.6, ior=1.4)
    wo = np.array([0.3, -0.2, -0.93])
    wo /= np.linalg.norm(wo)
    # (b) isotropic metal: the overwritten diffuse pick always lands on the specular lobe,
    #     so the sampled density is the specular component pdf alone
    mm = disney(metallic=1.0, roughness=0.55)
    wo2 = np.array([0.4, 0.1, 0.91])
    wo2 /= np.linalg.norm(wo2)
    for mat, w, comp in ((m, wo, lambda d: pdf(m, d, wo)),
                         (mm, wo2, lambda d: (pdf(mm, d, wo2) - 0.3 * max(d[2], 0) / math.pi) / 0.7)):
        dirs = []
        while len(dirs) < 30000:
            u, v, x = rng.random(3)
            s = po.bsdf(mat, 2, wo=w, u=u, v=v, w=x)
            if s[:3].any() and s[6] > 0:
                dirs.append(s[:3])
        chi2, dof = _chi2(np.array(dirs), comp)
        assert chi2 < dof + 6 * math.sqrt(2 * dof), (chi2, dof)
