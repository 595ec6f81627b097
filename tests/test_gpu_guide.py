"""Path-guiding network on the gfx950 matrix cores (no counterpart in Atrc).

* Inference (one-blob encoding, BF16 MFMA 16x16x32 GEMMs with FP32
  accumulation, ReLU, vMF-mixture decoding) against a numpy FP32 restatement
  that rounds features, weights and activations to BF16 where the kernel does;
  tolerance 2e-2 (BF16 has an 8-bit significand; the kernel's fast exp / sin /
  cos and its accumulation order differ from numpy's).
* Training gradient against torch autograd in FP32 on the CPU (rtol 2e-3), and
  one Adam step against its closed form.
"""
import math

import numpy as np
import pytest

from atrc_amd import capi

from conftest import log

pytestmark = pytest.mark.gpu

P = capi.GUIDE_PARAMS
LAYERS = [(256, 64), (64, 64), (64, 64), (64, 64), (64, 32)]


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for fi, fo in LAYERS:
        b = 1 / math.sqrt(fi)
        out.append(rng.uniform(-b, b, (fi, fo)).astype(np.float32).ravel())
        out.append(rng.uniform(-0.1, 0.1, fo).astype(np.float32))
    p = np.concatenate(out)
    assert p.size == P
    return p


def split(p):
    ws, o = [], 0
    for fi, fo in LAYERS:
        W = p[o:o + fi * fo].reshape(fi, fo); o += fi * fo
        b = p[o:o + fo]; o += fo
        ws.append((W, b))
    return ws


def bf16(x):
    """Round FP32 to BF16 (nearest even) and back."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def one_blob(x):
    x = np.clip(x, 0, 1)
    b = np.arange(32) + 0.5
    return np.exp(-0.5 * (x[:, :, None] * 32 - b) ** 2).reshape(len(x), 256)


def ref_infer(p, x):
    h = bf16(one_blob(x))
    for l, (W, b) in enumerate(split(p)):
        y = h.astype(np.float64) @ bf16(W).astype(np.float64) + b
        h = y if l == 4 else bf16(np.maximum(y, 0))
    o = h.reshape(len(x), 8, 4)
    th = math.pi / (1 + np.exp(-o[..., 1]))
    ph = 2 * math.pi / (1 + np.exp(-o[..., 2]))
    kap = np.exp(np.clip(o[..., 3], -4, 7))
    mu = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)
    w = np.exp(o[..., 0] - o[..., 0].max(1, keepdims=True))
    w /= w.sum(1, keepdims=True)
    return np.concatenate([np.concatenate([mu, kap[..., None]], -1).reshape(len(x), 32), w], 1)


def test_guide_inference_matches_reference(gpu_ctx):
    p = init_params(1) * 3   # larger weights: non-trivial activations
    rng = np.random.default_rng(2)
    for n in (1, 17, 1000, 70001, 300007):   # 300007 takes the 64-row-tile kernel (batches of >= 262144 rows)
        x = rng.random((n, 8)).astype(np.float32)
        log(f"[guide] inference n={n}")
        g = gpu_ctx.guide_infer(p, x)
        r = ref_infer(p, x)
        assert np.isfinite(g).all()
        np.testing.assert_allclose(g[:, 32:].sum(1), 1.0, rtol=1e-5)
        np.testing.assert_allclose(g, r, rtol=2e-2, atol=2e-2)


def torch_loss(p, samples, wscale):
    import torch

    pt = torch.tensor(p, dtype=torch.float32, requires_grad=True)
    x = torch.tensor(one_blob(samples[:, :8]), dtype=torch.float32)
    o, h = 0, x
    for l, (fi, fo) in enumerate(LAYERS):
        W = pt[o:o + fi * fo].reshape(fi, fo); o += fi * fo
        b = pt[o:o + fo]; o += fo
        h = h @ W + b
        if l < 4:
            h = torch.relu(h)
    out = h.reshape(-1, 8, 4)
    logw = torch.log_softmax(out[..., 0], dim=1)
    th = math.pi * torch.sigmoid(out[..., 1])
    ph = 2 * math.pi * torch.sigmoid(out[..., 2])
    kap = torch.exp(torch.clamp(out[..., 3], -4, 7))
    mu = torch.stack([torch.sin(th) * torch.cos(ph), torch.sin(th) * torch.sin(ph), torch.cos(th)], -1)
    d = torch.tensor(samples[:, 8:11], dtype=torch.float32)[:, None, :]
    logv = torch.log(kap) - math.log(2 * math.pi) - torch.log(-torch.expm1(-2 * kap)) + kap * ((mu * d).sum(-1) - 1)
    logq = torch.logsumexp(logw + logv, dim=1)
    w = torch.tensor(samples[:, 11], dtype=torch.float32)
    loss = -(w * wscale * logq).sum()
    loss.backward()
    return pt.grad.numpy()


def make_samples(rng, n):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    s = np.concatenate([rng.random((n, 8)), d, rng.exponential(1.0, (n, 1))], 1).astype(np.float32)
    return s


def test_guide_training_gradient_and_adam(gpu_ctx):
    rng = np.random.default_rng(3)
    p = init_params(4)
    s = make_samples(rng, 300)
    wscale = float(1 / s[:, 11].sum())
    log("[guide] training gradient vs torch autograd, 300 samples")
    m = np.zeros(P, np.float32)
    v = np.zeros(P, np.float32)
    p2, m2, v2, g = gpu_ctx.guide_train(p, m, v, s, wscale, lr=1e-2, t=1)
    ref = torch_loss(p, s, wscale)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(g, ref, rtol=2e-3, atol=2e-4 * scale)
    # Adam, t = 1: m = 0.1 g, v = 0.01 g^2, step = lr * g / (|g| + eps)
    np.testing.assert_allclose(m2, 0.1 * g, rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(v2, 0.01 * g * g, rtol=1e-5, atol=1e-20)
    np.testing.assert_allclose(p2, p - 1e-2 * g / (np.abs(g) + 1e-8), rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("n", [1000, 40000])
def test_guide_matrix_core_gradient(gpu_ctx, n):
    """The matrix-core training path (BF16 operands, FP32 accumulation) that
    renders use gives the autograd gradient to BF16 accuracy: per layer, the
    relative error of the weight and bias blocks is within 8 % and the
    direction within 4 degrees (BF16 rounding of activations, weights and
    output gradients compounds through the five layers; a wrong index or a
    dropped term gives errors of order 1).  1000 samples span several
    16-sample tiles, blocks and a partial last tile; 40,000 samples over the
    debug entry's 64 blocks of 256 samples make blocks take up to three
    iterations, whose gradients add to the block's partial."""
    rng = np.random.default_rng(11)
    p = init_params(7)
    s = make_samples(rng, n)
    wscale = float(1 / s[:, 11].sum())
    log(f"[guide] matrix-core training gradient vs torch autograd, {n} samples")
    g = gpu_ctx.guide_grad_mfma(p, s, wscale)
    ref = torch_loss(p, s, wscale)
    off = 0
    for l, (fi, fo) in enumerate([(256, 64), (64, 64), (64, 64), (64, 64), (64, 32)]):
        for name, size in (("W", fi * fo), ("b", fo)):
            a, b = g[off:off + size].astype(np.float64), ref[off:off + size]
            off += size
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            cosang = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300))
            log(f"  layer {l} {name}: relative error {rel:.4f}, cos {cosang:.6f}")
            assert rel < 0.08 and cosang > math.cos(math.radians(4.0)), (l, name)
    assert off == P


def test_guide_training_learns_a_direction(gpu_ctx):
    """A few hundred Adam steps on samples whose directions cluster around one
    axis concentrate the mixture there (the density at the axis grows)."""
    rng = np.random.default_rng(5)
    p = init_params(6)
    m = np.zeros(P, np.float32)
    v = np.zeros(P, np.float32)
    axis = np.array([0.3, -0.5, 0.81]); axis /= np.linalg.norm(axis)
    def density(pp, dirs, x):
        o = gpu_ctx.guide_infer(pp, x)
        mu = o[:, :32].reshape(-1, 8, 4)
        w = o[:, 32:]
        k = mu[..., 3]
        cosv = (mu[..., :3] * dirs[:, None, :]).sum(-1)
        vmf = k / (2 * math.pi * (1 - np.exp(-2 * k))) * np.exp(k * (cosv - 1))
        return (w * vmf).sum(1)
    x = rng.random((256, 8)).astype(np.float32)
    q0 = density(p, np.tile(axis, (256, 1)), x).mean()
    for t in range(1, 201):
        d = axis + 0.15 * rng.normal(size=(512, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        s = np.concatenate([rng.random((512, 8)), d, np.ones((512, 1))], 1).astype(np.float32)
        p, m, v, _ = gpu_ctx.guide_train(p, m, v, s, 1 / 512, lr=1e-2, t=t)
    q1 = density(p, np.tile(axis, (256, 1)), x).mean()
    log(f"  density at the axis: {q0:.3f} -> {q1:.3f}")
    assert q1 > 10 * q0
