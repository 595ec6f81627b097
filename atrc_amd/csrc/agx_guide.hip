// Path-guiding MLP on the gfx950 matrix cores: weight packing, batched
// inference (one-blob encoding, 4 x 64 ReLU, mixture decoding), training
// (forward, backward, KL-divergence gradient of the vMF mixture) and Adam.
// Layout constants are in agx_guide.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "agx_guide.h"
#include "agx_guide_kernels.h"

namespace agx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr float PI_G = 3.14159265358979323846f;

// v_mfma_f32_16x16x32_bf16 operand layouts (wave64):
//   A (16 x 32): lane l holds row l % 16, k = 8 * (l / 16) + i, i = 0..7
//   B (32 x 16): lane l holds column l % 16, k = 8 * (l / 16) + i
//   D (16 x 16): lane l holds column l % 16, rows 4 * (l / 16) + j, j = 0..3
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// LDS of one wave must be visible to its other lanes before they read it
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// one-blob feature of input value v for bin b: Gaussian of one bin width
__device__ __forceinline__ float one_blob(float v, int b) {
    const float x = v * float(G_BINS) - (float(b) + 0.5f);
    return __expf(-0.5f * x * x);
}

// Input k of a 32-wide k-step held as element i by lane group g (lanes 16 g ..
// 16 g + 15) of a fragment.  Layer 0 takes its inputs in natural order (the
// one-blob features are generated in registers in any order).  Layers 1..4
// take them in the order an MFMA output leaves them: output tile 2 s' (rows
// 0..15 of the k-step) gives lane group g rows 4 g .. 4 g + 3, output tile
// 2 s' + 1 (rows 16..31) rows 16 + 4 g .. 16 + 4 g + 3, so a layer's output
// fragments are the next layer's input fragments without any data movement.
__host__ __device__ constexpr int g_kperm(int l, int g, int i) {
    return l == 0 ? 8 * g + i : (i < 4 ? 4 * g + i : 16 + 4 * g + (i - 4));
}

// FP32 master parameters -> BF16 fragment-ordered weights + FP32 biases.
// packed[off(l) + ((s * T + t) * 64 + lane) * 8 + i] = W_l[32 s + g_kperm(l, lane / 16, i)][16 t + lane % 16]:
// the B fragment of x W (rows x inputs times inputs x outputs) and, identically,
// the A fragment of W^T x^T (outputs x inputs times inputs x rows).
__global__ void __launch_bounds__(256) k_guide_pack(const float *params, __bf16 *packed, float *bias) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < G_PACK) {
        int l = 0;
        while (l + 1 < G_LAYERS && gid >= g_pack_off(l + 1)) ++l;
        const int local = gid - g_pack_off(l);
        const int i = local & 7, lane = (local >> 3) & 63, frag = local >> 9;
        const int T = g_out(l) / 16;
        const int s = frag / T, t = frag % T;
        const int k = 32 * s + g_kperm(l, lane / 16, i), n = 16 * t + (lane % 16);
        packed[gid] = __bf16(params[g_w_off(l) + k * g_out(l) + n]);
    }
    if (gid < G_PACK_BWD) {
        // packed[G_PACK + off_bwd(l) + ((m * S + s) * 64 + lane) * 8 + i] =
        //     W_l[16 m + lane % 16][32 s + g_kperm(l, lane / 16, i)], S = outputs / 32
        int l = 1;
        while (l + 1 < G_LAYERS && gid >= g_pack_bwd_off(l + 1)) ++l;
        const int local = gid - g_pack_bwd_off(l);
        const int i = local & 7, lane = (local >> 3) & 63, frag = local >> 9;
        const int S = g_out(l) / 32;
        const int m = frag / S, sk = frag % S;
        const int k = 16 * m + lane % 16, n = 32 * sk + g_kperm(l, lane / 16, i);
        packed[G_PACK + gid] = __bf16(params[g_w_off(l) + k * g_out(l) + n]);
    }
    if (gid < G_BIASES) {
        int l = 0;
        while (l + 1 < G_LAYERS && gid >= g_bias_off(l + 1)) ++l;
        bias[gid] = params[g_b_off(l) + gid - g_bias_off(l)];
    }
}

// Batched inference: row i = in[8 i .. 8 i + 7] -> out[40 i ..]: 8 lobes of
// (mu.xyz, kappa) then 8 normalised lobe weights.
//
// Each layer is computed transposed, Y^T = W^T X^T: the packed weights are the
// A operand (16 outputs x 32 inputs), the activations the B operand (32 inputs
// x 16 rows), and an output fragment (16 outputs x 16 rows: lane l holds row
// l % 16, outputs 4 (l / 16) .. + 3) becomes, after bias, ReLU and BF16
// rounding in registers, the input fragment of the next layer in the packed
// input order (g_kperm).  Activations never leave registers.  A wave works on
// IR 16-row sub-tiles at once, so every weight fragment read from LDS feeds IR
// MFMAs.  Persistent blocks of IW waves; all BF16 weights and the biases stay
// in LDS for the block's lifetime (61 KB: two blocks per CU, so blocks of 8
// waves give each SIMD 4 waves).

constexpr int IW = 8;  // waves per inference block

struct InferSmem {
    __bf16 w[G_PACK];
    float b[G_BIASES];
    float x[IW][64 * G_IN];  // each wave's current tile inputs (at most 64 rows), row-major
};

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr float LOG2E_G = 1.44269504088896340736f;
constexpr float INV_E_G = 0.36787944117144233f;

// (lo, hi) rounded to a BF16 pair (nearest even, one v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t bf16_pair(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
}
// ReLU of a BF16 pair: a set sign bit reads as a negative int16, and max with 0
// clears it (rounding then ReLU equals ReLU then rounding)
__device__ __forceinline__ uint32_t relu_pair(float lo, float hi) {
    const s16x2 v = __builtin_bit_cast(s16x2, __builtin_convertvector((f32x2){lo, hi}, bf16x2));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(v, (s16x2){0, 0}));
}

// One-blob bins 8 g .. 8 g + 7 of x as a B fragment.  With d_i = 32 x - (8 g +
// i) - 1/2, bin i is exp(-d_i^2 / 2); neighbours follow by ratios:
// bin i + 1 = bin i * exp(d_i - 1/2) and bin i - 1 = bin i * exp(-d_i - 1/2),
// and each further step multiplies the ratio by 1/e.  The recurrence starts at
// bin 4, so the bins near the Gaussian's centre carry the rounding of an
// exponential of a small argument, as a direct evaluation would: three
// exponentials per eight bins instead of eight.
__device__ __forceinline__ bf16x8 one_blob8(float x, int g) {
    const float d = x * float(G_BINS) - (float(8 * g) + 4.5f);
    float v[8];
    v[4] = __builtin_amdgcn_exp2f(d * d * (-0.5f * LOG2E_G));
    float up = __builtin_amdgcn_exp2f((d - 0.5f) * LOG2E_G);
    float dn = __builtin_amdgcn_exp2f((-d - 0.5f) * LOG2E_G);
    v[5] = v[4] * up;
    up = up * INV_E_G;
    v[6] = v[5] * up;
    up = up * INV_E_G;
    v[7] = v[6] * up;
    v[3] = v[4] * dn;
    dn = dn * INV_E_G;
    v[2] = v[3] * dn;
    dn = dn * INV_E_G;
    v[1] = v[2] * dn;
    dn = dn * INV_E_G;
    v[0] = v[1] * dn;
    const u32x4 u = {bf16_pair(v[0], v[1]), bf16_pair(v[2], v[3]), bf16_pair(v[4], v[5]), bf16_pair(v[6], v[7])};
    return __builtin_bit_cast(bf16x8, u);
}

__device__ __forceinline__ float sigmoid_fast(float x) { return __builtin_amdgcn_rcpf(1.0f + __expf(-x)); }

// Forward weight fragment (A operand of W^T x^T) of layer l, k-step s, output tile t.
__device__ __forceinline__ bf16x8 wfrag(const __bf16 *w, int l, int s, int t, int lane) {
    return *reinterpret_cast<const bf16x8 *>(&w[g_pack_off(l) + ((s * (g_out(l) / 16) + t) * 64 + lane) * 8]);
}

// ReLU and BF16 rounding of a 64-output layer's fragments (bias already in
// the accumulators) -> the next layer's two k-step fragments of each sub-tile
template <int IR>
__device__ __forceinline__ void relu_frags(const f32x4 (&acc)[IR][4], bf16x8 (&h)[IR][2]) {
    #pragma unroll
    for (int q = 0; q < IR; ++q) {
        #pragma unroll
        for (int s = 0; s < 2; ++s) {
            const f32x4 a = acc[q][2 * s], b = acc[q][2 * s + 1];
            const u32x4 u = {relu_pair(a[0], a[1]), relu_pair(a[2], a[3]), relu_pair(b[0], b[1]), relu_pair(b[2], b[3])};
            h[q][s] = __builtin_bit_cast(bf16x8, u);
        }
    }
}

// 64 -> T * 16 outputs for the IR sub-tiles; the first k-step's accumulator
// input is the bias (lane group g: outputs 16 t + 4 g .. + 3)
template <int L, int IR>
__device__ __forceinline__ void chained_layer(const __bf16 *W, const float *B, int lane, const bf16x8 (&h)[IR][2],
                                              f32x4 (&acc)[IR][4]) {
    constexpr int T = g_out(L) / 16;
    f32x4 b[T];
    #pragma unroll
    for (int t = 0; t < T; ++t) b[t] = *reinterpret_cast<const f32x4 *>(&B[g_bias_off(L) + 16 * t + 4 * (lane / 16)]);
    #pragma unroll
    for (int s = 0; s < 2; ++s) {
        #pragma unroll
        for (int t = 0; t < T; ++t) {
            const bf16x8 w = wfrag(W, L, s, t, lane);
            #pragma unroll
            for (int q = 0; q < IR; ++q) acc[q][t] = mfma(w, h[q][s], s == 0 ? b[t] : acc[q][t]);
        }
    }
}

template <int IR>  // 16-row sub-tiles per wave
__global__ void __launch_bounds__(IW * 64, 4) k_guide_infer(const __bf16 *packed, const float *bias, const float *in,
                                                     float *out, const uint32_t *n_ptr, uint32_t n_fixed) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    InferSmem &sm = *reinterpret_cast<InferSmem *>(smem_raw);
    const uint32_t n = n_ptr ? *n_ptr : n_fixed;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    constexpr uint32_t ROWS = 16u * IR;  // rows of one wave tile
    if (uint32_t(blockIdx.x) * IW * ROWS >= n) return;  // whole block past the rows (block-uniform)
    {
        const uint4 *src = reinterpret_cast<const uint4 *>(packed);
        uint4 *dst = reinterpret_cast<uint4 *>(sm.w);
        for (int i = tid; i < G_PACK / 8; i += IW * 64) dst[i] = src[i];
        for (int i = tid; i < G_BIASES; i += IW * 64) sm.b[i] = bias[i];
    }
    __syncthreads();
    const int r = lane % 16, g = lane / 16;
    for (uint32_t tile = blockIdx.x * IW + uint32_t(wave); tile * ROWS < n; tile += gridDim.x * IW) {
        // wz is zero, laundered once per tile: the weights in LDS do not change over
        // the tile loop, and without it the compiler hoists every fragment load out
        // of the loop and spills the 240 registers they take
        int wz = 0;
        asm volatile("" : "+s"(wz));
        const __bf16 *W = sm.w + wz;
        const float *B = sm.b + wz;
        // layer 0: k-step s covers the 32 bins of input s; lane group g makes bins 8 g .. 8 g + 7;
        // the first k-step's accumulator input is the bias
        f32x4 acc[IR][4];
        {
            f32x4 b[4];
            #pragma unroll
            for (int t = 0; t < 4; ++t) b[t] = *reinterpret_cast<const f32x4 *>(&B[g_bias_off(0) + 16 * t + 4 * g]);
            #pragma unroll
            for (int q = 0; q < IR; ++q)
                #pragma unroll
                for (int t = 0; t < 4; ++t) acc[q][t] = b[t];
        }
        // the tile's inputs: 16 * IR rows of 8 floats, staged in LDS with 16-byte loads
        float *xs = sm.x[wave];
        wave_sync();  // the previous tile's reads of xs are done
        #pragma unroll
        for (int v = lane; v < 16 * IR * G_IN / 4; v += 64) {
            const uint32_t row = tile * ROWS + uint32_t(v / 2);
            const float4 a = row < n ? *reinterpret_cast<const float4 *>(in + size_t(row) * G_IN + 4 * (v % 2))
                                     : make_float4(0, 0, 0, 0);
            *reinterpret_cast<float4 *>(xs + 4 * v) = a;
        }
        wave_sync();
        #pragma unroll 2
        for (int s = 0; s < G_IN; ++s) {
            float x[IR];
            #pragma unroll
            for (int q = 0; q < IR; ++q) x[q] = fminf(1.0f, fmaxf(0.0f, xs[(16 * q + r) * G_IN + s]));
            bf16x8 w[4];
            #pragma unroll
            for (int t = 0; t < 4; ++t) w[t] = wfrag(W, 0, s, t, lane);
            #pragma unroll
            for (int q = 0; q < IR; ++q) {
                const bf16x8 f = one_blob8(x[q], g);
                #pragma unroll
                for (int t = 0; t < 4; ++t) acc[q][t] = mfma(w[t], f, acc[q][t]);
            }
        }
        bf16x8 h[IR][2];
        relu_frags<IR>(acc, h);
        chained_layer<1, IR>(W, B, lane, h, acc);
        relu_frags<IR>(acc, h);
        chained_layer<2, IR>(W, B, lane, h, acc);
        relu_frags<IR>(acc, h);
        chained_layer<3, IR>(W, B, lane, h, acc);
        relu_frags<IR>(acc, h);
        chained_layer<4, IR>(W, B, lane, h, acc);
        // decode: output tile t of lane group g holds lobe 4 t + g of row r, components
        // (weight logit, theta, phi, log kappa)
        #pragma unroll
        for (int q = 0; q < IR; ++q) {
            const uint32_t row = tile * ROWS + 16u * q + uint32_t(r);
            const bool valid = row < n;
            float logit[2];
            #pragma unroll
            for (int t = 0; t < 2; ++t) {
                const f32x4 o = acc[q][t];
                logit[t] = o[0];
                const float th = PI_G * sigmoid_fast(o[1]), ph = 2.0f * PI_G * sigmoid_fast(o[2]);
                const float kappa = __expf(fminf(G_LOGK_MAX, fmaxf(G_LOGK_MIN, o[3])));
                float st, ct, sp, cp;
                __sincosf(th, &st, &ct);
                __sincosf(ph, &sp, &cp);
                if (valid)
                    *reinterpret_cast<float4 *>(out + size_t(row) * G_ROW_OUT + 4 * (4 * t + g)) =
                        make_float4(st * cp, st * sp, ct, kappa);
            }
            // softmax of the 8 weight logits over the row's four lane groups
            float mx = fmaxf(logit[0], logit[1]);
            mx = fmaxf(mx, __shfl_xor(mx, 16));
            mx = fmaxf(mx, __shfl_xor(mx, 32));
            const float e0 = __expf(logit[0] - mx), e1 = __expf(logit[1] - mx);
            float sum = e0 + e1;
            sum += __shfl_xor(sum, 16);
            sum += __shfl_xor(sum, 32);
            const float inv = __builtin_amdgcn_rcpf(sum);
            if (valid) {
                out[size_t(row) * G_ROW_OUT + 4 * G_LOBES + g] = e0 * inv;
                out[size_t(row) * G_ROW_OUT + 4 * G_LOBES + 4 + g] = e1 * inv;
            }
        }
    }
}

// ---------------------------------------------------------------- training
// A training sample: the 8 network inputs, the sampled direction (world space,
// unit) and its weight F / p (incident radiance luminance over the density the
// direction was drawn with).  Loss = -sum_s w_s log q(dir_s) / sum_s w_s (the
// KL divergence between the normalised target and the mixture, up to a constant).
constexpr int G_SAMPLE = G_IN + 4;   // floats per sample
constexpr int TS = 16;               // samples per chunk

struct TrainSmem {
    float x0[TS][G_FEAT];            // one-blob features
    float h[4][TS][G_HID];           // hidden activations (after ReLU)
    float o[TS][G_OUT];              // raw outputs, then d loss / d outputs
    float d[2][TS][G_HID];           // back-propagated gradients (ping-pong)
    float smp[TS][G_SAMPLE];
};

// log of the vMF density kappa / (2 pi (1 - e^-2kappa)) e^{kappa (mu.w - 1)}
__device__ __forceinline__ float vmf_log_norm(float kappa) {
    return logf(kappa) - logf(2.0f * PI_G) - logf(-expm1f(-2.0f * kappa));
}

// d(-log q(dir)) / d(raw outputs) of one sample, times its weight
__device__ void mixture_grad(const float *o, const float *dir, float w, float *g) {
    float logit_max = -3.0e38f;
    for (int j = 0; j < G_LOBES; ++j) logit_max = fmaxf(logit_max, o[4 * j]);
    float wsum = 0, wj[G_LOBES];
    for (int j = 0; j < G_LOBES; ++j) { wj[j] = expf(o[4 * j] - logit_max); wsum += wj[j]; }
    float lt[G_LOBES], kap[G_LOBES], cosm[G_LOBES], s1[G_LOBES], s2[G_LOBES], mdt[G_LOBES], mdp[G_LOBES];
    float lmax = -3.0e38f;
    for (int j = 0; j < G_LOBES; ++j) {
        wj[j] /= wsum;
        s1[j] = 1.0f / (1.0f + expf(-o[4 * j + 1]));
        s2[j] = 1.0f / (1.0f + expf(-o[4 * j + 2]));
        const float th = PI_G * s1[j], ph = 2.0f * PI_G * s2[j];
        const float lk = fminf(G_LOGK_MAX, fmaxf(G_LOGK_MIN, o[4 * j + 3]));
        kap[j] = expf(lk);
        const float st = sinf(th), ct = cosf(th), sp = sinf(ph), cp = cosf(ph);
        cosm[j] = st * cp * dir[0] + st * sp * dir[1] + ct * dir[2];
        mdt[j] = ct * cp * dir[0] + ct * sp * dir[1] - st * dir[2];   // dir . dmu/dtheta
        mdp[j] = -st * sp * dir[0] + st * cp * dir[1];                 // dir . dmu/dphi
        lt[j] = logf(fmaxf(wj[j], 1e-30f)) + vmf_log_norm(kap[j]) + kap[j] * (cosm[j] - 1.0f);
        lmax = fmaxf(lmax, lt[j]);
    }
    float qs = 0;
    for (int j = 0; j < G_LOBES; ++j) qs += expf(lt[j] - lmax);
    const float logq = lmax + logf(qs);
    for (int j = 0; j < G_LOBES; ++j) {
        const float r = expf(lt[j] - logq);
        g[4 * j] = -w * (r - wj[j]);
        g[4 * j + 1] = -w * r * kap[j] * mdt[j] * PI_G * s1[j] * (1.0f - s1[j]);
        g[4 * j + 2] = -w * r * kap[j] * mdp[j] * 2.0f * PI_G * s2[j] * (1.0f - s2[j]);
        const float lk_raw = o[4 * j + 3];
        const bool free_k = lk_raw > G_LOGK_MIN && lk_raw < G_LOGK_MAX;
        const float e2 = expf(-2.0f * kap[j]);
        const float dlogv = 1.0f / kap[j] - 2.0f * e2 / (-expm1f(-2.0f * kap[j])) + (cosm[j] - 1.0f);
        g[4 * j + 3] = free_k ? -w * r * dlogv * kap[j] : 0.0f;
    }
}

// Forward + backward of chunks of TS samples; block b accumulates its
// gradient contributions into partial[b][G_PARAMS] (no atomics: the sum over
// blocks is taken by k_guide_adam in a fixed order, so training is deterministic).
__global__ void __launch_bounds__(256) k_guide_train(const float *params, const float *samples,
                                                     const uint32_t *n_dev, uint32_t cap, const float *wsum_dev,
                                                     float *partial) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    TrainSmem &sm = *reinterpret_cast<TrainSmem *>(smem_raw);
    const uint32_t n = min(*n_dev, cap);
    const float wsum = *wsum_dev;
    const float wscale = wsum > 0.0f ? 1.0f / wsum : 0.0f;
    float *grad = partial + size_t(blockIdx.x) * G_PARAMS;
    const int tid = threadIdx.x;
    for (int i = tid; i < G_PARAMS; i += 256) grad[i] = 0.0f;
    for (uint32_t c0 = blockIdx.x * TS; c0 < n; c0 += gridDim.x * TS) {
        const int cnt = int(min(uint32_t(TS), n - c0));
        __syncthreads();
        for (int i = tid; i < TS * G_SAMPLE; i += 256) {
            const int sI = i / G_SAMPLE, k = i % G_SAMPLE;
            sm.smp[sI][k] = sI < cnt ? samples[size_t(c0 + sI) * G_SAMPLE + k] : 0.0f;
        }
        __syncthreads();
        for (int i = tid; i < TS * G_FEAT; i += 256) {
            const int sI = i / G_FEAT, f = i % G_FEAT;
            const float v = fminf(1.0f, fmaxf(0.0f, sm.smp[sI][f / G_BINS]));
            sm.x0[sI][f] = one_blob(v, f % G_BINS);
        }
        __syncthreads();
        // forward: thread (n = tid % 64, quarter = tid / 64) computes 4 samples of column n
        for (int l = 0; l < G_LAYERS; ++l) {
            const int in_l = g_in(l), out_l = g_out(l);
            const float *W = params + g_w_off(l), *B = params + g_b_off(l);
            for (int job = tid; job < out_l * (TS / 4); job += 256) {
                const int col = job % out_l, sq = job / out_l;
                float acc[4];
                for (int q = 0; q < 4; ++q) acc[q] = B[col];
                for (int k = 0; k < in_l; ++k) {
                    const float wk = W[k * out_l + col];
                    for (int q = 0; q < 4; ++q) {
                        const int sI = 4 * sq + q;
                        const float a = l == 0 ? sm.x0[sI][k] : sm.h[l - 1][sI][k];
                        acc[q] += a * wk;
                    }
                }
                for (int q = 0; q < 4; ++q) {
                    const int sI = 4 * sq + q;
                    if (l < G_LAYERS - 1) sm.h[l][sI][col] = fmaxf(acc[q], 0.0f);
                    else sm.o[sI][col] = acc[q];
                }
            }
            __syncthreads();
        }
        // d loss / d outputs
        if (tid < TS) {
            float g[G_OUT];
            const float w = tid < cnt ? sm.smp[tid][G_IN + 3] * wscale : 0.0f;
            if (w != 0.0f) mixture_grad(sm.o[tid], &sm.smp[tid][G_IN], w, g);
            else for (int j = 0; j < G_OUT; ++j) g[j] = 0.0f;
            for (int j = 0; j < G_OUT; ++j) sm.o[tid][j] = g[j];
        }
        __syncthreads();
        // backward: layer l has input activations a (x0 or h[l-1]) and output gradient dy
        for (int l = G_LAYERS - 1; l >= 0; --l) {
            const int in_l = g_in(l), out_l = g_out(l);
            const float *W = params + g_w_off(l);
            float *dy_buf = l == G_LAYERS - 1 ? &sm.o[0][0] : &sm.d[l % 2][0][0];
            const int dy_stride = l == G_LAYERS - 1 ? G_OUT : G_HID;
            // weight and bias gradients: thread owns (k, col) pairs
            for (int job = tid; job < in_l * out_l; job += 256) {
                const int k = job / out_l, col = job % out_l;
                float acc = 0;
                for (int sI = 0; sI < TS; ++sI) {
                    const float a = l == 0 ? sm.x0[sI][k] : sm.h[l - 1][sI][k];
                    acc += a * dy_buf[sI * dy_stride + col];
                }
                grad[g_w_off(l) + job] += acc;
            }
            for (int col = tid; col < out_l; col += 256) {
                float acc = 0;
                for (int sI = 0; sI < TS; ++sI) acc += dy_buf[sI * dy_stride + col];
                grad[g_b_off(l) + col] += acc;
            }
            if (l > 0) {
                // gradient wrt the input activations of layer l, through ReLU
                float *dx = &sm.d[(l - 1) % 2][0][0];
                for (int job = tid; job < TS * in_l; job += 256) {
                    const int sI = job / in_l, k = job % in_l;
                    float acc = 0;
                    for (int col = 0; col < out_l; ++col) acc += dy_buf[sI * dy_stride + col] * W[k * out_l + col];
                    dx[sI * G_HID + k] = sm.h[l - 1][sI][k] > 0.0f ? acc : 0.0f;
                }
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------- matrix-core training
// One Adam step over the sample ring is k_gtrain_step, then the fixed-order
// reduction over blocks and Adam, with no host wait.  A k_gtrain_step block
// takes TW * 16 samples per iteration, one 16-sample tile per wave:
//   - the forward pass through all five layers and the back-propagation
//     dY_5 -> dY_1, activations and gradients kept in registers (the inference
//     kernel's chaining; the weights packed again as the A operand of W_l dY^T);
//   - per layer l, from the top: every wave writes its H_l and dY_(l+1) tiles
//     to LDS, feature-major over the block's samples, and the block takes
//     dW_l = H_l^T dY_(l+1), one 16 x 16 tile per wave with K = the block's
//     samples, and db_l (MFMAs with an all-ones A operand); layer 0's one-blob
//     features are rebuilt in LDS 64 at a time;
//   - each tile goes to the block's own partial gradient (stored on the
//     block's first iteration, added by the same lane on later ones), so no
//     activation leaves the CU and the sums keep a fixed order.
// dY_(l+1) = d loss / d (H_l W_l + b_l); dY_l = (dY_(l+1) W_l^T) masked by
// ReLU'(H_l), H_0 = the one-blob features.
constexpr int TW = 16;               // waves per k_gtrain_step block
constexpr int TB_S = TW * 16;        // samples per block iteration
constexpr int TS_ROW = TB_S + 8;       // an LDS row: 16 bytes of padding put the 16 rows an operand reads in distinct banks

struct StepSmem {
    __bf16 w[G_PACK_ALL];  // forward fragments, then the back-propagation fragments
    float b[G_BIASES];
    __bf16 a[G_HID][TS_ROW];  // H_l, or 64 one-blob features, of the block's samples
    union {
        __bf16 d[G_HID][TS_ROW];       // dY_(l+1) of the block's samples
        float smp[TW][16 * G_SAMPLE];  // during the forward pass: each wave's tile of samples, row-major
    } u;
};

// Back-propagation weight fragment (A operand of W_l dY^T): inputs 16 m .., k-step s over outputs.
__device__ __forceinline__ bf16x8 wbfrag(const __bf16 *w, int l, int m, int s, int lane) {
    return *reinterpret_cast<const bf16x8 *>(&w[G_PACK + g_pack_bwd_off(l) + ((m * (g_out(l) / 32) + s) * 64 + lane) * 8]);
}

// A wave's tile of 32 * NF features held as B fragments (sample lane % 16,
// features in g_kperm order) -> dst[feature][16 wave + sample].  Lanes of
// neighbouring samples r and r ^ 1 swap halves (a DPP quad permute), so each
// lane writes whole dwords: the even lane feature 2 d's pair of samples, the
// odd lane feature 2 d + 1's, for the d-th dword of each fragment.
template <int NF>
__device__ __forceinline__ void stage_tile(__bf16 (*dst)[TS_ROW], const bf16x8 (&h)[2], int wave, int lane) {
    const int r = lane % 16, g = lane / 16, odd = r & 1;
    __bf16 *base = &dst[0][16 * wave + (r & ~1)];
    #pragma unroll
    for (int s = 0; s < NF; ++s) {
        const u32x4 u = __builtin_bit_cast(u32x4, h[s]);
        #pragma unroll
        for (int d = 0; d < 4; ++d) {
            const uint32_t mine = u[d];
            // quad_perm [1, 0, 3, 2]: the value of lane r ^ 1
            const uint32_t other = uint32_t(__builtin_amdgcn_update_dpp(0, int(mine), 0xB1, 0xF, 0xF, false));
            const uint32_t word = odd ? (mine & 0xFFFF0000u) | (other >> 16) : (other << 16) | (mine & 0xFFFFu);
            const int i = 2 * d + odd;
            *reinterpret_cast<uint32_t *>(base + (32 * s + g_kperm(1, g, i)) * TS_ROW) = word;
        }
    }
}

// d(-w log q(dir)) / d(raw outputs) with the 8 lobes of a row spread over its
// four lanes (lane group g holds lobes g and 4 + g: o[t] = outputs 4 (4 t + g)
// .. + 3); the sums and maxima over lobes are taken across the four lanes.
// The per-lobe terms are mixture_grad's, with the hardware exponential,
// logarithm, sine, cosine and reciprocal, and log kappa taken from the clamped
// output itself.
__device__ void mixture_grad_lanes(const f32x4 (&o)[2], const float *dir, float w, float (&gr)[2][4]) {
    float lm = fmaxf(o[0][0], o[1][0]);
    lm = fmaxf(lm, __shfl_xor(lm, 16));
    lm = fmaxf(lm, __shfl_xor(lm, 32));
    float wj[2];
    wj[0] = __expf(o[0][0] - lm);
    wj[1] = __expf(o[1][0] - lm);
    float ws = wj[0] + wj[1];
    ws += __shfl_xor(ws, 16);
    ws += __shfl_xor(ws, 32);
    const float inv_ws = __builtin_amdgcn_rcpf(ws);
    float lt[2], kap[2], cosm[2], s1[2], s2[2], mdt[2], mdp[2], om[2];
    #pragma unroll
    for (int t = 0; t < 2; ++t) {
        wj[t] = wj[t] * inv_ws;
        s1[t] = sigmoid_fast(o[t][1]);
        s2[t] = sigmoid_fast(o[t][2]);
        const float th = PI_G * s1[t], ph = 2.0f * PI_G * s2[t];
        const float lk = fminf(G_LOGK_MAX, fmaxf(G_LOGK_MIN, o[t][3]));
        kap[t] = __expf(lk);
        float st, ct, sp, cp;
        __sincosf(th, &st, &ct);
        __sincosf(ph, &sp, &cp);
        cosm[t] = st * cp * dir[0] + st * sp * dir[1] + ct * dir[2];
        mdt[t] = ct * cp * dir[0] + ct * sp * dir[1] - st * dir[2];
        mdp[t] = -st * sp * dir[0] + st * cp * dir[1];
        om[t] = 1.0f - __expf(-2.0f * kap[t]);  // 1 - e^(-2 kappa) >= 0.035 for kappa >= e^-4
        lt[t] = __logf(fmaxf(wj[t], 1e-30f)) + (lk - 1.8378770664093453f - __logf(om[t])) + kap[t] * (cosm[t] - 1.0f);
    }
    float lmax = fmaxf(lt[0], lt[1]);
    lmax = fmaxf(lmax, __shfl_xor(lmax, 16));
    lmax = fmaxf(lmax, __shfl_xor(lmax, 32));
    float qs = __expf(lt[0] - lmax) + __expf(lt[1] - lmax);
    qs += __shfl_xor(qs, 16);
    qs += __shfl_xor(qs, 32);
    const float logq = lmax + __logf(qs);
    #pragma unroll
    for (int t = 0; t < 2; ++t) {
        const float r = __expf(lt[t] - logq);
        gr[t][0] = -w * (r - wj[t]);
        gr[t][1] = -w * r * kap[t] * mdt[t] * PI_G * s1[t] * (1.0f - s1[t]);
        gr[t][2] = -w * r * kap[t] * mdp[t] * 2.0f * PI_G * s2[t] * (1.0f - s2[t]);
        const float lk_raw = o[t][3];
        const bool free_k = lk_raw > G_LOGK_MIN && lk_raw < G_LOGK_MAX;
        const float dlogv = __builtin_amdgcn_rcpf(kap[t]) - 2.0f * (1.0f - om[t]) * __builtin_amdgcn_rcpf(om[t]) +
                            (cosm[t] - 1.0f);
        gr[t][3] = free_k ? -w * r * dlogv * kap[t] : 0.0f;
    }
}

// dY_l^T = W_l dY_(l+1)^T masked by ReLU'(H_l): dy_in = dY_(l+1) as B fragments
// (K = outputs of layer l in g_kperm order), h = H_l as B fragments; out = dY_l
// as B fragments of the layer below.
template <int L>
__device__ __forceinline__ void bwd_layer(const __bf16 *W, int lane, const bf16x8 (&dy_in)[2], const bf16x8 (&h)[2],
                                          bf16x8 (&out)[2]) {
    constexpr int S = g_out(L) / 32;
    f32x4 acc[4];
    #pragma unroll
    for (int m = 0; m < 4; ++m) {
        acc[m] = f32x4{0, 0, 0, 0};
        #pragma unroll
        for (int s = 0; s < S; ++s) acc[m] = mfma(wbfrag(W, L, m, s, lane), dy_in[s], acc[m]);
    }
    // output tile m holds inputs 16 m + 4 g + j: element (m % 2) * 4 + j of h[m / 2]
    #pragma unroll
    for (int s = 0; s < 2; ++s) {
        float v[8];
        #pragma unroll
        for (int j = 0; j < 4; ++j) {
            v[j] = float(h[s][j]) > 0.0f ? acc[2 * s][j] : 0.0f;
            v[4 + j] = float(h[s][4 + j]) > 0.0f ? acc[2 * s + 1][j] : 0.0f;
        }
        const u32x4 u = {bf16_pair(v[0], v[1]), bf16_pair(v[2], v[3]), bf16_pair(v[4], v[5]), bf16_pair(v[6], v[7])};
        out[s] = __builtin_bit_cast(bf16x8, u);
    }
}

// acc = A[16 mt ..][samples] x B[16 nt ..][samples]^T over the block's TB_S
// samples; with ONES the A operand is all ones (every row: B's column sums)
template <bool ONES>
__device__ __forceinline__ f32x4 block_tile(const __bf16 (*A)[TS_ROW], int mt, const __bf16 (*Bm)[TS_ROW], int nt,
                                            int lane) {
    const int r = lane % 16, g = lane / 16;
    bf16x8 ones;
    #pragma unroll
    for (int i = 0; i < 8; ++i) ones[i] = __bf16(1.0f);
    f32x4 acc[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    #pragma unroll 2
    for (int ks = 0; ks < TB_S / 32; ++ks) {
        const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&Bm[16 * nt + r][32 * ks + 8 * g]);
        const bf16x8 a = ONES ? ones : *reinterpret_cast<const bf16x8 *>(&A[16 * mt + r][32 * ks + 8 * g]);
        acc[ks % 2] = mfma(a, b, acc[ks % 2]);
    }
    return acc[0] + acc[1];
}

// D layout (lane: column lane % 16, rows 4 (lane / 16) + j) -> rows 16 mt + ..,
// columns 16 nt + .. of layer l's weight gradient in the block's partial
__device__ __forceinline__ void put_w(float *part, int l, int mt, int nt, int lane, f32x4 v, bool first) {
    const int r = lane % 16, g = lane / 16;
    #pragma unroll
    for (int j = 0; j < 4; ++j) {
        float *q = part + g_w_off(l) + (16 * mt + 4 * g + j) * g_out(l) + 16 * nt + r;
        *q = first ? v[j] : *q + v[j];
    }
}
// row 0 of a column-sum tile -> biases 16 nt .. 16 nt + 15 of layer l
__device__ __forceinline__ void put_b(float *part, int l, int nt, int lane, f32x4 v, bool first) {
    if (lane < 16) {
        float *q = part + g_b_off(l) + 16 * nt + lane;
        *q = first ? v[0] : *q + v[0];
    }
}

// Layer L (1..3) over the block's samples: H_L and dY_(L+1) to LDS, the
// back-propagation dY_L into out, then dW_L (wave: rows 16 (wave / 4) ..,
// columns 16 (wave % 4) ..) and db_L (waves 4 L .. 4 L + 3)
template <int L>
__device__ __forceinline__ void train_layer64(StepSmem &sm, const __bf16 *W, const bf16x8 (&h)[2],
                                              const bf16x8 (&dy)[2], bf16x8 (&out)[2], float *part, int wave,
                                              int lane, bool first) {
    __syncthreads();  // the layer above has read sm.a and sm.u.d
    stage_tile<2>(sm.a, h, wave, lane);
    stage_tile<2>(sm.u.d, dy, wave, lane);
    bwd_layer<L>(W, lane, dy, h, out);
    __syncthreads();
    put_w(part, L, wave / 4, wave % 4, lane, block_tile<false>(sm.a, wave / 4, sm.u.d, wave % 4, lane), first);
    if (wave / 4 == L) put_b(part, L, wave % 4, lane, block_tile<true>(sm.a, 0, sm.u.d, wave % 4, lane), first);
}

__global__ void __launch_bounds__(TW * 64) k_gtrain_step(const __bf16 *packed, const float *bias,
                                                         const float *samples, const uint32_t *n_dev, uint32_t cap,
                                                         const float *wsum_dev, float *partial) {
    static_assert(TW == 16, "the dW tiling and the one-blob staging assume 16 waves");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    StepSmem &sm = *reinterpret_cast<StepSmem *>(smem_raw);
    const uint32_t n = min(*n_dev, cap);
    const uint32_t tiles = (n + 15) / 16;  // rows past n in the last tile get w = 0
    const float wsum = *wsum_dev;
    const float wscale = wsum > 0.0f ? 1.0f / wsum : 0.0f;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    float *part = partial + size_t(blockIdx.x) * G_PARAMS;
    if (uint32_t(blockIdx.x) * TW >= tiles) {  // no samples for this block (block-uniform): a zero partial
        for (int i = tid; i < G_PARAMS; i += TW * 64) part[i] = 0.0f;
        return;
    }
    {
        const uint4 *src = reinterpret_cast<const uint4 *>(packed);
        uint4 *dst = reinterpret_cast<uint4 *>(sm.w);
        for (int i = tid; i < G_PACK_ALL / 8; i += TW * 64) dst[i] = src[i];
        for (int i = tid; i < G_BIASES; i += TW * 64) sm.b[i] = bias[i];
    }
    const int r = lane % 16, g = lane / 16;
    // block-uniform trip count: every wave reaches every barrier
    for (uint32_t base = blockIdx.x * TW; base < tiles; base += gridDim.x * TW) {
        const bool first = base == blockIdx.x * TW;
        const uint32_t tile = base + uint32_t(wave);
        int wz = 0;
        asm volatile("" : "+s"(wz));
        const __bf16 *W = sm.w + wz;
        const float *B = sm.b + wz;
        const uint32_t row = tile * 16u + uint32_t(r);
        const bool valid = row < n;
        __syncthreads();  // the weights are in LDS; the previous iteration has read sm.a and sm.u
        // the tile's 16 samples (12 floats each), staged in LDS with 16-byte loads; rows past n read as zeros
        float *xs = sm.u.smp[wave];
        if (lane < 16 * G_SAMPLE / 4) {
            const uint32_t rr = tile * 16u + uint32_t(lane / (G_SAMPLE / 4));
            const float4 v = rr < n ? *reinterpret_cast<const float4 *>(samples + size_t(tile) * 16 * G_SAMPLE + 4 * lane)
                                    : make_float4(0, 0, 0, 0);
            *reinterpret_cast<float4 *>(xs + 4 * lane) = v;
        }
        wave_sync();
        const float *sx = xs + r * G_SAMPLE;
        // forward
        f32x4 acc[1][4];
        #pragma unroll
        for (int t = 0; t < 4; ++t) acc[0][t] = *reinterpret_cast<const f32x4 *>(&B[g_bias_off(0) + 16 * t + 4 * g]);
        #pragma unroll 2
        for (int s = 0; s < G_IN; ++s) {
            const float x = fminf(1.0f, fmaxf(0.0f, sx[s]));
            const bf16x8 f = one_blob8(x, g);
            #pragma unroll
            for (int t = 0; t < 4; ++t) acc[0][t] = mfma(wfrag(W, 0, s, t, lane), f, acc[0][t]);
        }
        bf16x8 h1[1][2], h2[1][2], h3[1][2], h4[1][2];
        relu_frags<1>(acc, h1);
        chained_layer<1, 1>(W, B, lane, h1, acc);
        relu_frags<1>(acc, h2);
        chained_layer<2, 1>(W, B, lane, h2, acc);
        relu_frags<1>(acc, h3);
        chained_layer<3, 1>(W, B, lane, h3, acc);
        relu_frags<1>(acc, h4);
        chained_layer<4, 1>(W, B, lane, h4, acc);
        // d loss / d outputs: lobes g and 4 + g of row r
        float dir[3] = {0.0f, 0.0f, 1.0f};
        float w = 0.0f;
        if (valid) {
            dir[0] = sx[G_IN]; dir[1] = sx[G_IN + 1]; dir[2] = sx[G_IN + 2];
            w = sx[G_IN + 3] * wscale;
        }
        const f32x4 o[2] = {acc[0][0], acc[0][1]};
        float gr[2][4];
        mixture_grad_lanes(o, dir, w, gr);
        if (w == 0.0f)
            #pragma unroll
            for (int t = 0; t < 2; ++t)
                #pragma unroll
                for (int c = 0; c < 4; ++c) gr[t][c] = 0.0f;
        bf16x8 dy5[2];
        {
            const u32x4 u = {bf16_pair(gr[0][0], gr[0][1]), bf16_pair(gr[0][2], gr[0][3]), bf16_pair(gr[1][0], gr[1][1]),
                             bf16_pair(gr[1][2], gr[1][3])};
            dy5[0] = __builtin_bit_cast(bf16x8, u);
            dy5[1] = dy5[0];  // unused: the output layer has one k-step
        }
        // layer 4 (64 -> 32): dW_4 on waves 0..7, db_4 on waves 8 and 9
        __syncthreads();  // every wave has read its samples: sm.u holds dY from here on
        stage_tile<2>(sm.a, h4[0], wave, lane);
        stage_tile<1>(sm.u.d, dy5, wave, lane);
        bf16x8 d4[2];
        bwd_layer<4>(W, lane, dy5, h4[0], d4);
        __syncthreads();
        if (wave < 8)
            put_w(part, 4, wave / 2, wave % 2, lane, block_tile<false>(sm.a, wave / 2, sm.u.d, wave % 2, lane), first);
        else if (wave < 10)
            put_b(part, 4, wave - 8, lane, block_tile<true>(sm.a, 0, sm.u.d, wave - 8, lane), first);
        bf16x8 d3[2], d2[2], d1[2];
        train_layer64<3>(sm, W, h3[0], d4, d3, part, wave, lane, first);
        train_layer64<2>(sm, W, h2[0], d3, d2, part, wave, lane, first);
        train_layer64<1>(sm, W, h1[0], d2, d1, part, wave, lane, first);
        // layer 0 (256 -> 64): dY_1 stays in LDS while the one-blob features
        // of inputs 2 p and 2 p + 1 (64 features) take sm.a in pass p
        __syncthreads();  // layer 1 has read sm.a and sm.u.d
        stage_tile<2>(sm.u.d, d1, wave, lane);
        // thread: samples 2 sp and 2 sp + 1 of the block, bins 8 part8 .. + 7 of input 2 p + ii
        const int sp = tid % (TB_S / 2), part8 = (tid / (TB_S / 2)) % 4, ii = tid / (2 * TB_S);
        for (int p = 0; p < G_IN / 2; ++p) {
            float x[2];
            #pragma unroll
            for (int q = 0; q < 2; ++q) {
                const uint32_t rw = base * 16u + uint32_t(2 * sp + q);
                x[q] = rw < n ? fminf(1.0f, fmaxf(0.0f, samples[size_t(rw) * G_SAMPLE + 2 * p + ii])) : 0.0f;
            }
            if (p) __syncthreads();  // pass p - 1 has read sm.a (the loads above are in flight meanwhile)
            {
                const bf16x8 v0 = one_blob8(x[0], part8), v1 = one_blob8(x[1], part8);
                #pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const bf16x2 pr = {v0[k], v1[k]};
                    *reinterpret_cast<bf16x2 *>(&sm.a[32 * ii + 8 * part8 + k][2 * sp]) = pr;
                }
            }
            __syncthreads();
            put_w(part, 0, 4 * p + wave / 4, wave % 4, lane, block_tile<false>(sm.a, wave / 4, sm.u.d, wave % 4, lane),
                  first);
            if (p == 0 && wave < 4) put_b(part, 0, wave, lane, block_tile<true>(sm.a, 0, sm.u.d, wave, lane), first);
        }
    }
}

// one block: advance the ring, then a fixed-order sum of the valid weights
__global__ void __launch_bounds__(1024) k_guide_ring_commit(uint32_t *ring, const uint32_t *added, uint32_t cap,
                                                            const float *samples, float *wsum) {
    __shared__ float part[1024];
    __shared__ uint32_t valid_s;
    if (threadIdx.x == 0) {
        const uint32_t a = min(*added, cap);
        ring[0] = (ring[0] + a) % cap;
        valid_s = min(ring[1] + a, cap);
        ring[1] = valid_s;
    }
    __syncthreads();
    const uint32_t valid = valid_s;
    // eight independent loads in flight per thread, added in a fixed order
    float a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t i = threadIdx.x; i < valid; i += 8 * 1024) {
        #pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint32_t k = i + uint32_t(u) * 1024u;
            if (k < valid) a[u] += samples[size_t(k) * G_SAMPLE + G_IN + 3];
        }
    }
    part[threadIdx.x] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *wsum = part[0];
}

// The fixed-order sum runs in two passes so that every CU takes part: pass 1
// adds each group of `per` consecutive block partials into the group's first
// partial (GR_GROUPS x ceil(params / 256) blocks of independent loads), pass 2
// adds the group sums in group order.
constexpr int GR_GROUPS = 16;

__global__ void __launch_bounds__(256) k_guide_reduce_groups(float *partial, int blocks, int per) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int b0 = int(blockIdx.y) * per;
    if (i >= G_PARAMS || b0 >= blocks) return;
    const int b1 = min(blocks, b0 + per);
    float g = partial[size_t(b0) * G_PARAMS + i];
    for (int b = b0 + 1; b < b1; ++b) g += partial[size_t(b) * G_PARAMS + i];
    partial[size_t(b0) * G_PARAMS + i] = g;
}

__global__ void __launch_bounds__(256) k_guide_reduce(const float *partial, int blocks, int per, float *grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G_PARAMS) return;
    float g = 0;
    for (int b = 0; b < blocks; b += per) g += partial[size_t(b) * G_PARAMS + i];
    grad[i] = g;
}

// Adam step of parameter i (k_guide_adam's update and repacking)
__device__ __forceinline__ void adam_update(int i, float g, float *params, float *m, float *v, float lr, float b1,
                                            float b2, float eps, int t, __bf16 *packed, float *bias);

// The fixed-order reduction and Adam in one launch: a block takes RA_COLS
// parameters; its thread (j, y) adds the partials of group y (blocks y per ..
// y per + per - 1, in order) of parameter j, and the group sums are then
// added in group order: the sums and their order are k_guide_reduce_groups'
// and k_guide_reduce's, with no pass through memory in between.
constexpr int RA_COLS = 64;

__global__ void __launch_bounds__(RA_COLS * GR_GROUPS) k_guide_reduce_adam(const float *partial, int blocks, int per,
                                                                           float *params, float *m, float *v, float lr,
                                                                           float b1, float b2, float eps, int t,
                                                                           __bf16 *packed, float *bias) {
    __shared__ float gs[GR_GROUPS][RA_COLS];
    const int j = threadIdx.x % RA_COLS, y = threadIdx.x / RA_COLS;
    const int i = blockIdx.x * RA_COLS + j;
    const int b0 = y * per;
    float g = 0.0f;
    if (i < G_PARAMS && b0 < blocks) {
        const int b1g = min(blocks, b0 + per);
        g = partial[size_t(b0) * G_PARAMS + i];
        for (int b = b0 + 1; b < b1g; ++b) g += partial[size_t(b) * G_PARAMS + i];
    }
    gs[y][j] = g;
    __syncthreads();
    if (y != 0 || i >= G_PARAMS) return;
    float sum = 0.0f;
    for (int yy = 0; yy < GR_GROUPS && yy * per < blocks; ++yy) sum += gs[yy][j];
    adam_update(i, sum, params, m, v, lr, b1, b2, eps, t, packed, bias);
}

// One Adam step on the FP32 master parameters: m, v moments; lr; bias
// corrections for step t.  When packed is given, the updated parameter is
// also written where k_guide_pack puts it (BF16 weight in fragment order, or
// FP32 bias), so the next forward pass needs no separate packing launch.
__device__ __forceinline__ void adam_update(int i, float g, float *params, float *m, float *v, float lr, float b1,
                                            float b2, float eps, int t, __bf16 *packed, float *bias) {
    const float mi = b1 * m[i] + (1 - b1) * g;
    const float vi = b2 * v[i] + (1 - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1 - powf(b1, float(t)));
    const float vh = vi / (1 - powf(b2, float(t)));
    const float p = params[i] - lr * mh / (sqrtf(vh) + eps);
    params[i] = p;
    if (!packed) return;
    int l = 0;
    while (l + 1 < G_LAYERS && i >= g_w_off(l + 1)) ++l;
    const int local = i - g_w_off(l);
    if (local >= g_in(l) * g_out(l)) {  // bias
        bias[g_bias_off(l) + local - g_in(l) * g_out(l)] = p;
        return;
    }
    // inverse of k_guide_pack's index map (g_kperm)
    const int k = local / g_out(l), n = local % g_out(l), T = g_out(l) / 16;
    const int sI = k / 32, tI = n / 16, kk = k % 32;
    const int grp = l == 0 ? kk / 8 : (kk % 16) / 4;
    const int elem = l == 0 ? kk % 8 : (kk < 16 ? 0 : 4) + kk % 4;
    packed[g_pack_off(l) + ((sI * T + tI) * 64 + grp * 16 + n % 16) * 8 + elem] = __bf16(p);
    if (l > 0) {  // the back-propagation fragments (k_guide_pack's second map)
        const int m = k / 16, sk = n / 32, nn = n % 32;
        const int gb = (nn % 16) / 4, eb = (nn < 16 ? 0 : 4) + nn % 4;
        packed[G_PACK + g_pack_bwd_off(l) + ((m * (g_out(l) / 32) + sk) * 64 + gb * 16 + k % 16) * 8 + eb] = __bf16(p);
    }
}

__global__ void __launch_bounds__(256) k_guide_adam(float *params, float *m, float *v, const float *grad, float lr,
                                                    float b1, float b2, float eps, int t, __bf16 *packed,
                                                    float *bias) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G_PARAMS) return;
    adam_update(i, grad[i], params, m, v, lr, b1, b2, eps, t, packed, bias);
}

}  // namespace

size_t guide_smem_bytes() { return sizeof(InferSmem) > sizeof(StepSmem) ? sizeof(InferSmem) : sizeof(StepSmem); }
size_t guide_train_smem_bytes() { return sizeof(TrainSmem); }

void launch_guide_train(const float *params, const float *samples, const uint32_t *n_dev, uint32_t cap,
                        const float *wsum_dev, float *partial, int blocks, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_train, dim3(blocks), dim3(256), sizeof(TrainSmem), st, params, samples, n_dev, cap,
                       wsum_dev, partial);
}

void launch_guide_train_mfma(const __bf16 *packed, const float *bias, const float *samples, const uint32_t *n_dev,
                             uint32_t cap, const float *wsum_dev, float *partial, int blocks, hipStream_t st) {
    hipLaunchKernelGGL(k_gtrain_step, dim3(blocks), dim3(TW * 64), sizeof(StepSmem), st, packed, bias, samples, n_dev,
                       cap, wsum_dev, partial);
}

void launch_guide_ring_commit(uint32_t *ring, const uint32_t *added, uint32_t cap, const float *samples,
                              float *wsum, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_ring_commit, dim3(1), dim3(1024), 0, st, ring, added, cap, samples, wsum);
}

void launch_guide_reduce(float *partial, int blocks, float *grad, hipStream_t st) {
    const int per = (blocks + GR_GROUPS - 1) / GR_GROUPS;
    hipLaunchKernelGGL(k_guide_reduce_groups, dim3((G_PARAMS + 255) / 256, GR_GROUPS), dim3(256), 0, st, partial,
                       blocks, per);
    hipLaunchKernelGGL(k_guide_reduce, dim3((G_PARAMS + 255) / 256), dim3(256), 0, st, partial, blocks, per, grad);
}

void launch_guide_reduce_adam(const float *partial, int blocks, float *params, float *m, float *v, float lr, float b1,
                              float b2, float eps, int t, hipStream_t st, __bf16 *packed, float *bias) {
    const int per = (blocks + GR_GROUPS - 1) / GR_GROUPS;
    hipLaunchKernelGGL(k_guide_reduce_adam, dim3((G_PARAMS + RA_COLS - 1) / RA_COLS), dim3(RA_COLS * GR_GROUPS), 0, st,
                       partial, blocks, per, params, m, v, lr, b1, b2, eps, t, packed, bias);
}

void launch_guide_adam(float *params, float *m, float *v, const float *grad, float lr, float b1, float b2, float eps,
                       int t, hipStream_t st, __bf16 *packed, float *bias) {
    hipLaunchKernelGGL(k_guide_adam, dim3((G_PARAMS + 255) / 256), dim3(256), 0, st, params, m, v, grad, lr, b1, b2,
                       eps, t, packed, bias);
}

void launch_guide_pack(const float *params, __bf16 *packed, float *bias, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_pack, dim3((G_PACK + 255) / 256), dim3(256), 0, st, params, packed, bias);
}

hipError_t guide_prepare(int device) {
    int lds = 0;
    hipError_t e = hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
    if (e != hipSuccess) return e;
    if (size_t(lds) < sizeof(InferSmem) || size_t(lds) < sizeof(TrainSmem) || size_t(lds) < sizeof(StepSmem))
        return hipErrorInvalidConfiguration;
    for (const void *k : {reinterpret_cast<const void *>(k_guide_infer<1>), reinterpret_cast<const void *>(k_guide_infer<4>)}) {
        e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(sizeof(InferSmem)));
        if (e != hipSuccess) return e;
    }
    e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_guide_train),
                            hipFuncAttributeMaxDynamicSharedMemorySize, int(sizeof(TrainSmem)));
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(k_gtrain_step),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(sizeof(StepSmem)));
}

void launch_guide_infer(const __bf16 *packed, const float *bias, const float *in, float *out, const uint32_t *n_dev,
                        uint32_t n_max, hipStream_t st) {
    if (!n_max) return;
    // persistent: 2 blocks per CU resident (LDS), enough to cover the rows.  Waves
    // take 64-row tiles (each weight fragment feeds 4 MFMAs) once every CU gets
    // at least one such tile per wave; smaller batches take 16-row tiles.
    if (n_max >= 512u * IW * 64u) {
        const uint32_t blocks = (n_max + IW * 64 - 1) / (IW * 64);
        hipLaunchKernelGGL(k_guide_infer<4>, dim3(blocks < 512u ? blocks : 512u), dim3(IW * 64), sizeof(InferSmem), st,
                           packed, bias, in, out, n_dev, n_max);
    } else {
        const uint32_t blocks = (n_max + IW * 16 - 1) / (IW * 16);
        hipLaunchKernelGGL(k_guide_infer<1>, dim3(blocks < 512u ? blocks : 512u), dim3(IW * 64), sizeof(InferSmem), st,
                           packed, bias, in, out, n_dev, n_max);
    }
}

}  // namespace agx
