// Path-guiding network (no counterpart in Atrc): an online-trained MLP that
// maps a surface vertex (position, normal, view direction) to a mixture of
// von Mises-Fisher lobes over incident directions; the renderer samples the
// continuation direction from a one-sample mixture of the BSDF and this
// distribution.
//
// Network: one-blob encoding of 8 inputs in [0, 1] (32 Gaussian bins each,
// sigma = one bin) -> 256 features -> 4 hidden layers of 64 (ReLU) -> 32
// outputs = 8 lobes x (weight logit, theta, phi, log kappa).  Weights and
// activations are BF16 in the matrix-core GEMMs, accumulation FP32; the master
// weights and the Adam state are FP32.
#pragma once

#include <cstddef>
#include <cstdint>

namespace agx {

constexpr int G_IN = 8;
constexpr int G_BINS = 32;
constexpr int G_FEAT = G_IN * G_BINS;  // 256
constexpr int G_HID = 64;
constexpr int G_OUT = 32;
constexpr int G_LOBES = 8;
constexpr int G_LAYERS = 5;
constexpr int G_ROW_OUT = 5 * G_LOBES;  // decoded row: 8 x (mu.xyz, kappa), then 8 weights

// layer l maps g_in(l) -> g_out(l); FP32 master parameters: W_l [in][out] row-major then b_l [out]
__host__ __device__ constexpr int g_in(int l) { return l == 0 ? G_FEAT : G_HID; }
__host__ __device__ constexpr int g_out(int l) { return l == G_LAYERS - 1 ? G_OUT : G_HID; }
__host__ __device__ constexpr int g_w_off(int l) {
    int o = 0;
    for (int k = 0; k < l; ++k) o += g_in(k) * g_out(k) + g_out(k);
    return o;
}
__host__ __device__ constexpr int g_b_off(int l) { return g_w_off(l) + g_in(l) * g_out(l); }
constexpr int G_PARAMS = g_w_off(G_LAYERS);               // 31008
// BF16 weights packed in matrix-core fragment order, layer after layer
__host__ __device__ constexpr int g_pack_off(int l) {
    int o = 0;
    for (int k = 0; k < l; ++k) o += g_in(k) * g_out(k);
    return o;
}
constexpr int G_PACK = g_pack_off(G_LAYERS);              // 30720 bf16
// BF16 weights of layers 1..4 packed again as the A operand of back-propagation
// (W_l rows = the layer's inputs), after the forward fragments
__host__ __device__ constexpr int g_pack_bwd_off(int l) {
    int o = 0;
    for (int k = 1; k < l; ++k) o += g_in(k) * g_out(k);
    return o;
}
constexpr int G_PACK_BWD = g_pack_bwd_off(G_LAYERS);     // 14336 bf16
constexpr int G_PACK_ALL = G_PACK + G_PACK_BWD;
__host__ __device__ constexpr int g_bias_off(int l) {
    int o = 0;
    for (int k = 0; k < l; ++k) o += g_out(k);
    return o;
}
constexpr int G_BIASES = g_bias_off(G_LAYERS);            // 288

// output decoding (kept in one place for inference, training and tests)
constexpr float G_LOGK_MIN = -4.0f, G_LOGK_MAX = 7.0f;

}  // namespace agx
