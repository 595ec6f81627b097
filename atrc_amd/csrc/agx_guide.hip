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
constexpr int ACT_STRIDE = G_HID + 8;      // bf16 per staged activation row (16-byte aligned rows)
constexpr int OST_STRIDE = G_OUT + 1;      // floats per staged output row

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

__device__ __forceinline__ float sigmoid(float x) { return 1.0f / (1.0f + __expf(-x)); }

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
    if (gid < G_BIASES) {
        int l = 0;
        while (l + 1 < G_LAYERS && gid >= g_bias_off(l + 1)) ++l;
        bias[gid] = params[g_b_off(l) + gid - g_bias_off(l)];
    }
}

struct GuideSmem {
    __bf16 w[G_PACK];
    float b[G_BIASES];
    __bf16 act[4][16][ACT_STRIDE];
    float ost[4][16][OST_STRIDE];
};

// Hidden layer l (64 -> 64 or 64 -> 32) for the 16 rows staged in act[wave]:
// reads the activations, then (after the wave has read them) overwrites them
// with ReLU(x W + b), or writes the raw outputs to ost for the last layer.
// The A fragment (16 rows x 32 inputs) of k-step s from a staged activation
// row, in the packed weights' input order (g_kperm).
__device__ __forceinline__ bf16x8 act_frag(const __bf16 *row, int s, int g) {
    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
    const bf16x4 lo = *reinterpret_cast<const bf16x4 *>(row + 32 * s + 4 * g);
    const bf16x4 hi = *reinterpret_cast<const bf16x4 *>(row + 32 * s + 16 + 4 * g);
    return bf16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

template <int L>
__device__ __forceinline__ void hidden_layer(GuideSmem &sm, int wave, int lane) {
    constexpr int T = g_out(L) / 16;
    const int r = lane % 16, g = lane / 16;
    const bf16x8 a0 = act_frag(&sm.act[wave][r][0], 0, g);
    const bf16x8 a1 = act_frag(&sm.act[wave][r][0], 1, g);
    f32x4 acc[T];
    #pragma unroll
    for (int t = 0; t < T; ++t) {
        acc[t] = f32x4{0, 0, 0, 0};
        const bf16x8 *w0 = reinterpret_cast<const bf16x8 *>(&sm.w[g_pack_off(L) + ((0 * T + t) * 64 + lane) * 8]);
        const bf16x8 *w1 = reinterpret_cast<const bf16x8 *>(&sm.w[g_pack_off(L) + ((1 * T + t) * 64 + lane) * 8]);
        acc[t] = mfma(a0, *w0, acc[t]);
        acc[t] = mfma(a1, *w1, acc[t]);
    }
    wave_sync();
    #pragma unroll
    for (int t = 0; t < T; ++t) {
        const int col = 16 * t + r;
        const float bias = sm.b[g_bias_off(L) + col];
        #pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float v = acc[t][j] + bias;
            if (L == G_LAYERS - 1) sm.ost[wave][4 * g + j][col] = v;
            else sm.act[wave][4 * g + j][col] = __bf16(v > 0.0f ? v : 0.0f);
        }
    }
    wave_sync();
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
};

// wz is zero, laundered once per tile: the weights in LDS do not change over
// the tile loop, and without it the compiler hoists every fragment load out of
// the loop and spills the 240 registers they take
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

__device__ __forceinline__ bf16x8 wfrag(const InferSmem &sm, int l, int s, int t, int lane, int wz) {
    return *reinterpret_cast<const bf16x8 *>(&sm.w[wz + g_pack_off(l) + ((s * (g_out(l) / 16) + t) * 64 + lane) * 8]);
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
__device__ __forceinline__ void chained_layer(const InferSmem &sm, int lane, int wz, const bf16x8 (&h)[IR][2],
                                              f32x4 (&acc)[IR][4]) {
    constexpr int T = g_out(L) / 16;
    f32x4 b[T];
    #pragma unroll
    for (int t = 0; t < T; ++t) b[t] = *reinterpret_cast<const f32x4 *>(&sm.b[wz + g_bias_off(L) + 16 * t + 4 * (lane / 16)]);
    #pragma unroll
    for (int s = 0; s < 2; ++s) {
        #pragma unroll
        for (int t = 0; t < T; ++t) {
            const bf16x8 w = wfrag(sm, L, s, t, lane, wz);
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
        int wz = 0;
        asm volatile("" : "+s"(wz));
        // layer 0: k-step s covers the 32 bins of input s; lane group g makes bins 8 g .. 8 g + 7;
        // the first k-step's accumulator input is the bias
        f32x4 acc[IR][4];
        {
            f32x4 b[4];
            #pragma unroll
            for (int t = 0; t < 4; ++t) b[t] = *reinterpret_cast<const f32x4 *>(&sm.b[wz + g_bias_off(0) + 16 * t + 4 * g]);
            #pragma unroll
            for (int q = 0; q < IR; ++q)
                #pragma unroll
                for (int t = 0; t < 4; ++t) acc[q][t] = b[t];
        }
        #pragma unroll 2
        for (int s = 0; s < G_IN; ++s) {
            float x[IR];
            #pragma unroll
            for (int q = 0; q < IR; ++q) {
                const uint32_t row = tile * ROWS + 16u * q + uint32_t(r);
                x[q] = row < n ? fminf(1.0f, fmaxf(0.0f, in[size_t(row) * G_IN + s])) : 0.0f;
            }
            bf16x8 w[4];
            #pragma unroll
            for (int t = 0; t < 4; ++t) w[t] = wfrag(sm, 0, s, t, lane, wz);
            #pragma unroll
            for (int q = 0; q < IR; ++q) {
                const bf16x8 f = one_blob8(x[q], g);
                #pragma unroll
                for (int t = 0; t < 4; ++t) acc[q][t] = mfma(w[t], f, acc[q][t]);
            }
        }
        bf16x8 h[IR][2];
        relu_frags<IR>(acc, h);
        chained_layer<1, IR>(sm, lane, wz, h, acc);
        relu_frags<IR>(acc, h);
        chained_layer<2, IR>(sm, lane, wz, h, acc);
        relu_frags<IR>(acc, h);
        chained_layer<3, IR>(sm, lane, wz, h, acc);
        relu_frags<IR>(acc, h);
        chained_layer<4, IR>(sm, lane, wz, h, acc);
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
// The same loss and gradient as k_guide_train, layer by layer on the matrix
// cores: BF16 operands (weights, activations, output gradients), FP32
// accumulation and FP32 master weights.  One step =
//   k_gtrain_forward: forward of every sample (activations H1..H4 to HBM) and
//     d loss / d outputs (dY5);
//   k_gtrain_backward<L>, L = 4..0: dW_L = H_L^T dY_{L+1} and db_L = sum dY_{L+1}
//     per block over its samples (partial[block]), and dY_L = (dY_{L+1} W_L^T)
//     masked by ReLU'(H_L) for the layer below;
//   then the fixed-order reduction over blocks and Adam, as before.
// Activations and gradients: bf16 [sample][64] rows (dY5 uses 32 of them).
constexpr int GT_CHUNK = 32;                     // samples per MFMA k-step in dW
constexpr int GT_PAD = 8;                        // bf16 padding of LDS rows

__global__ void __launch_bounds__(256) k_gtrain_forward(const __bf16 *packed, const float *bias, const float *samples,
                                                        const uint32_t *n_dev, uint32_t cap, const float *wsum_dev,
                                                        __bf16 *acts, __bf16 *dy5) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    GuideSmem &sm = *reinterpret_cast<GuideSmem *>(smem_raw);
    const uint32_t n = min(*n_dev, cap);
    const float wsum = *wsum_dev;
    const float wscale = wsum > 0.0f ? 1.0f / wsum : 0.0f;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    if (uint32_t(blockIdx.x) * 64u >= n) return;
    {
        const uint4 *src = reinterpret_cast<const uint4 *>(packed);
        uint4 *dst = reinterpret_cast<uint4 *>(sm.w);
        for (int i = tid; i < G_PACK / 8; i += 256) dst[i] = src[i];
        for (int i = tid; i < G_BIASES; i += 256) sm.b[i] = bias[i];
    }
    __syncthreads();
    const int r = lane % 16, g = lane / 16;
    for (uint32_t tile = blockIdx.x * 4u + uint32_t(wave); tile * 16u < n; tile += gridDim.x * 4u) {
        const uint32_t row = tile * 16u + uint32_t(r);
        const bool valid = row < n;
        float x[G_IN];
        #pragma unroll
        for (int k = 0; k < G_IN; ++k) {
            const float v = valid ? samples[size_t(row) * G_SAMPLE + k] : 0.0f;
            x[k] = fminf(1.0f, fmaxf(0.0f, v));
        }
        f32x4 acc[4];
        #pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = f32x4{0, 0, 0, 0};
        #pragma unroll
        for (int k = 0; k < G_IN; ++k) {
            bf16x8 a;
            #pragma unroll
            for (int i = 0; i < 8; ++i) a[i] = __bf16(one_blob(x[k], 8 * g + i));
            #pragma unroll
            for (int t = 0; t < 4; ++t) {
                const bf16x8 *w = reinterpret_cast<const bf16x8 *>(&sm.w[g_pack_off(0) + ((k * 4 + t) * 64 + lane) * 8]);
                acc[t] = mfma(a, *w, acc[t]);
            }
        }
        #pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int col = 16 * t + r;
            const float b0 = sm.b[g_bias_off(0) + col];
            #pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float v = acc[t][j] + b0;
                sm.act[wave][4 * g + j][col] = __bf16(v > 0.0f ? v : 0.0f);
            }
        }
        wave_sync();
        // H_l (input of layer l) for l = 1..4 goes to HBM before layer l overwrites it
        auto store_act = [&](int l) {
            const uint32_t base = tile * 16u;
            for (int e = lane; e < 16 * G_HID / 8; e += 64) {
                const int rr = e / (G_HID / 8), c8 = e % (G_HID / 8);
                if (base + uint32_t(rr) < n)
                    *reinterpret_cast<uint4 *>(acts + (size_t(l - 1) * cap + base + rr) * G_HID + 8 * c8) =
                        *reinterpret_cast<const uint4 *>(&sm.act[wave][rr][8 * c8]);
            }
        };
        store_act(1);
        hidden_layer<1>(sm, wave, lane);
        store_act(2);
        hidden_layer<2>(sm, wave, lane);
        store_act(3);
        hidden_layer<3>(sm, wave, lane);
        store_act(4);
        hidden_layer<4>(sm, wave, lane);
        // d loss / d outputs of row r (one lane per row)
        if (g == 0 && valid) {
            const float *smp = samples + size_t(row) * G_SAMPLE;
            float gr[G_OUT];
            const float w = smp[G_IN + 3] * wscale;
            if (w != 0.0f) mixture_grad(&sm.ost[wave][r][0], smp + G_IN, w, gr);
            else for (int j = 0; j < G_OUT; ++j) gr[j] = 0.0f;
            for (int j = 0; j < G_OUT; ++j) dy5[size_t(row) * G_HID + j] = __bf16(gr[j]);
        }
        wave_sync();
    }
}

template <int L>
struct GtBwdSmem {
    static constexpr int IN = g_in(L), OUT = g_out(L);
    __bf16 w[IN][OUT + GT_PAD];                  // W_L row-major (input feature, output)
    __bf16 ht[IN][GT_CHUNK + GT_PAD];            // input activations, feature-major
    __bf16 hs[GT_CHUNK][(L > 0 ? IN : 8) + GT_PAD];  // input activations, sample-major (ReLU mask)
    __bf16 ds[GT_CHUNK][OUT + GT_PAD];           // dY_{L+1}, sample-major
    __bf16 dt[OUT][GT_CHUNK + GT_PAD];           // dY_{L+1}, feature-major
};

template <int L>
__global__ void __launch_bounds__(256) k_gtrain_backward(const float *params, const float *samples,
                                                         const uint32_t *n_dev, uint32_t cap, const __bf16 *acts,
                                                         const __bf16 *dy_in, __bf16 *dy_out, float *partial) {
    using S = GtBwdSmem<L>;
    constexpr int IN = S::IN, OUT = S::OUT;
    constexpr int MT = IN / 16, NT = OUT / 16;            // dW tiles
    constexpr int DW_PER_WAVE = (MT * NT) / 4;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    S &sm = *reinterpret_cast<S *>(smem_raw);
    const uint32_t n = min(*n_dev, cap);
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int r = lane % 16, g = lane / 16;
    for (int i = tid; i < IN * OUT; i += 256) sm.w[i / OUT][i % OUT] = __bf16(params[g_w_off(L) + i]);
    // this block's samples: [s_begin, s_end), whole chunks
    const uint32_t chunks = (n + GT_CHUNK - 1) / GT_CHUNK;
    const uint32_t per = (chunks + gridDim.x - 1) / gridDim.x;
    const uint32_t c_begin = blockIdx.x * per, c_end = min(chunks, c_begin + per);
    f32x4 dw[DW_PER_WAVE];
    #pragma unroll
    for (int t = 0; t < DW_PER_WAVE; ++t) dw[t] = f32x4{0, 0, 0, 0};
    float db = 0.0f;
    for (uint32_t c = c_begin; c < c_end; ++c) {
        const uint32_t s0 = c * GT_CHUNK;
        __syncthreads();
        // stage dY_{L+1} (both layouts) and the input activations
        for (int e = tid; e < GT_CHUNK * OUT; e += 256) {
            const int sI = e / OUT, o = e % OUT;
            const __bf16 v = s0 + sI < n ? dy_in[size_t(s0 + sI) * G_HID + o] : __bf16(0.0f);
            sm.ds[sI][o] = v;
            sm.dt[o][sI] = v;
        }
        if constexpr (L == 0) {
            for (int e = tid; e < GT_CHUNK * IN; e += 256) {
                const int sI = e / IN, f = e % IN;
                float v = 0.0f;
                if (s0 + sI < n) {
                    const float xin = fminf(1.0f, fmaxf(0.0f, samples[size_t(s0 + sI) * G_SAMPLE + f / G_BINS]));
                    v = one_blob(xin, f % G_BINS);
                }
                sm.ht[f][sI] = __bf16(v);
            }
        } else {
            for (int e = tid; e < GT_CHUNK * IN; e += 256) {
                const int sI = e / IN, f = e % IN;
                const __bf16 v = s0 + sI < n ? acts[(size_t(L - 1) * cap + s0 + sI) * G_HID + f] : __bf16(0.0f);
                sm.ht[f][sI] = v;
                sm.hs[sI][f] = v;
            }
        }
        __syncthreads();
        // bias gradient: thread o sums its column over the chunk
        if (tid < OUT)
            for (int sI = 0; sI < GT_CHUNK; ++sI) db += float(sm.ds[sI][tid]);
        // dW tiles of this wave: A = H^T (feature rows, sample k), B = dY (sample k, output cols)
        #pragma unroll
        for (int t = 0; t < DW_PER_WAVE; ++t) {
            const int tile = wave * DW_PER_WAVE + t, mt = tile / NT, nt = tile % NT;
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&sm.ht[16 * mt + r][8 * g]);
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&sm.dt[16 * nt + r][8 * g]);
            dw[t] = mfma(a, b, dw[t]);
        }
        if constexpr (L > 0) {
            // dY_L = dY_{L+1} W_L^T masked by ReLU'(H_L): 2 x MT tiles of (16 samples x 16 features)
            constexpr int DH_TILES = 2 * MT, DH_PER_WAVE = DH_TILES / 4;
            #pragma unroll
            for (int t = 0; t < DH_PER_WAVE; ++t) {
                const int tile = wave * DH_PER_WAVE + t, ms = tile / MT, ni = tile % MT;
                f32x4 acc = f32x4{0, 0, 0, 0};
                #pragma unroll
                for (int kk = 0; kk < OUT / 32; ++kk) {
                    const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&sm.ds[16 * ms + r][32 * kk + 8 * g]);
                    const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&sm.w[16 * ni + r][32 * kk + 8 * g]);
                    acc = mfma(a, b, acc);
                }
                const int f = 16 * ni + r;
                #pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int sI = 16 * ms + 4 * g + j;
                    if (s0 + sI < n) {
                        const float h = float(sm.hs[sI][f]);
                        dy_out[size_t(s0 + sI) * G_HID + f] = __bf16(h > 0.0f ? acc[j] : 0.0f);
                    }
                }
            }
        }
    }
    // this block's partial gradient of layer L (zeros when it had no samples)
    float *part = partial + size_t(blockIdx.x) * G_PARAMS;
    #pragma unroll
    for (int t = 0; t < DW_PER_WAVE; ++t) {
        const int tile = wave * DW_PER_WAVE + t, mt = tile / NT, nt = tile % NT;
        const int col = 16 * nt + r;
        #pragma unroll
        for (int j = 0; j < 4; ++j) part[g_w_off(L) + (16 * mt + 4 * g + j) * OUT + col] = dw[t][j];
    }
    if (tid < OUT) part[g_b_off(L) + tid] = db;
}

// Sum of the per-block partial gradients in a fixed order.
// one block: advance the ring, then a fixed-order sum of the valid weights
__global__ void __launch_bounds__(256) k_guide_ring_commit(uint32_t *ring, const uint32_t *added, uint32_t cap,
                                                           const float *samples, float *wsum) {
    __shared__ float part[256];
    __shared__ uint32_t valid_s;
    if (threadIdx.x == 0) {
        const uint32_t a = min(*added, cap);
        ring[0] = (ring[0] + a) % cap;
        valid_s = min(ring[1] + a, cap);
        ring[1] = valid_s;
    }
    __syncthreads();
    const uint32_t valid = valid_s;
    float acc = 0.0f;
    for (uint32_t i = threadIdx.x; i < valid; i += 256) acc += samples[size_t(i) * G_SAMPLE + G_IN + 3];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
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

// One Adam step on the FP32 master parameters: m, v moments; lr; bias
// corrections for step t.  When packed is given, the updated parameter is
// also written where k_guide_pack puts it (BF16 weight in fragment order, or
// FP32 bias), so the next forward pass needs no separate packing launch.
__global__ void __launch_bounds__(256) k_guide_adam(float *params, float *m, float *v, const float *grad, float lr,
                                                    float b1, float b2, float eps, int t, __bf16 *packed,
                                                    float *bias) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G_PARAMS) return;
    const float g = grad[i];
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
}

}  // namespace

size_t guide_smem_bytes() { return sizeof(InferSmem) > sizeof(GuideSmem) ? sizeof(InferSmem) : sizeof(GuideSmem); }
size_t guide_train_smem_bytes() { return sizeof(TrainSmem); }

void launch_guide_train(const float *params, const float *samples, const uint32_t *n_dev, uint32_t cap,
                        const float *wsum_dev, float *partial, int blocks, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_train, dim3(blocks), dim3(256), sizeof(TrainSmem), st, params, samples, n_dev, cap,
                       wsum_dev, partial);
}

void launch_guide_train_mfma(const float *params, const __bf16 *packed, const float *bias, const float *samples,
                             const uint32_t *n_dev, uint32_t cap, const float *wsum_dev, __bf16 *acts, __bf16 *dy,
                             float *partial, int blocks, hipStream_t st) {
    // forward: enough persistent blocks to cover the ring's 16-row tiles
    const uint32_t tiles = (cap + 63) / 64;
    hipLaunchKernelGGL(k_gtrain_forward, dim3(std::min<uint32_t>(tiles, 512u)), dim3(256), sizeof(GuideSmem), st,
                       packed, bias, samples, n_dev, cap, wsum_dev, acts, dy);
    __bf16 *dy_a = dy, *dy_b = dy + size_t(cap) * G_HID;
    hipLaunchKernelGGL(k_gtrain_backward<4>, dim3(blocks), dim3(256), sizeof(GtBwdSmem<4>), st, params, samples, n_dev,
                       cap, acts, dy_a, dy_b, partial);
    hipLaunchKernelGGL(k_gtrain_backward<3>, dim3(blocks), dim3(256), sizeof(GtBwdSmem<3>), st, params, samples, n_dev,
                       cap, acts, dy_b, dy_a, partial);
    hipLaunchKernelGGL(k_gtrain_backward<2>, dim3(blocks), dim3(256), sizeof(GtBwdSmem<2>), st, params, samples, n_dev,
                       cap, acts, dy_a, dy_b, partial);
    hipLaunchKernelGGL(k_gtrain_backward<1>, dim3(blocks), dim3(256), sizeof(GtBwdSmem<1>), st, params, samples, n_dev,
                       cap, acts, dy_b, dy_a, partial);
    hipLaunchKernelGGL(k_gtrain_backward<0>, dim3(blocks), dim3(256), sizeof(GtBwdSmem<0>), st, params, samples, n_dev,
                       cap, acts, dy_a, nullptr, partial);
}

void launch_guide_ring_commit(uint32_t *ring, const uint32_t *added, uint32_t cap, const float *samples,
                              float *wsum, hipStream_t st) {
    hipLaunchKernelGGL(k_guide_ring_commit, dim3(1), dim3(256), 0, st, ring, added, cap, samples, wsum);
}

void launch_guide_reduce(float *partial, int blocks, float *grad, hipStream_t st) {
    const int per = (blocks + GR_GROUPS - 1) / GR_GROUPS;
    hipLaunchKernelGGL(k_guide_reduce_groups, dim3((G_PARAMS + 255) / 256, GR_GROUPS), dim3(256), 0, st, partial,
                       blocks, per);
    hipLaunchKernelGGL(k_guide_reduce, dim3((G_PARAMS + 255) / 256), dim3(256), 0, st, partial, blocks, per, grad);
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
    if (size_t(lds) < sizeof(GuideSmem) || size_t(lds) < sizeof(TrainSmem)) return hipErrorInvalidConfiguration;
    e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_guide_infer<1>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, int(sizeof(InferSmem)));
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_guide_infer<4>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, int(sizeof(InferSmem)));
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_guide_train),
                            hipFuncAttributeMaxDynamicSharedMemorySize, int(sizeof(TrainSmem)));
    if (e != hipSuccess) return e;
    if (size_t(lds) < sizeof(GuideSmem) || size_t(lds) < sizeof(GtBwdSmem<0>)) return hipErrorInvalidConfiguration;
    e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_gtrain_forward),
                            hipFuncAttributeMaxDynamicSharedMemorySize, int(sizeof(GuideSmem)));
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute(reinterpret_cast<const void *>(k_gtrain_backward<0>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(sizeof(GtBwdSmem<0>)));
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
