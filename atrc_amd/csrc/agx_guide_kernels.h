// Launchers of the path-guiding network kernels (agx_guide.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace agx {

// dynamic LDS bytes of the inference kernel
size_t guide_smem_bytes();
// checks the device's LDS limit against the kernels' needs and raises their
// dynamic-LDS attribute; hipErrorInvalidConfiguration if the device has too little
hipError_t guide_prepare(int device);
// FP32 master parameters (G_PARAMS) -> BF16 fragment-ordered weights (G_PACK) + FP32 biases (G_BIASES)
void launch_guide_pack(const float *params, __bf16 *packed, float *bias, hipStream_t st);
// rows: n_dev (device count, may be NULL) or n_max; in: 8 floats per row; out: G_ROW_OUT floats per row
void launch_guide_infer(const __bf16 *packed, const float *bias, const float *in, float *out, const uint32_t *n_dev,
                        uint32_t n_max, hipStream_t st);

// training: min(*n_dev, cap) samples of G_IN + 4 floats (inputs, direction,
// weight); the loss weights are the samples' weights / *wsum_dev; blocks x
// G_PARAMS partial gradients
size_t guide_train_smem_bytes();
void launch_guide_train(const float *params, const float *samples, const uint32_t *n_dev, uint32_t cap,
                        const float *wsum_dev, float *partial, int blocks, hipStream_t st);
// sums the per-block partial gradients in a fixed order (overwrites partial)
void launch_guide_reduce(float *partial, int blocks, float *grad, hipStream_t st);
// the same gradient on the matrix cores (BF16 operands, FP32 accumulation),
// one fused kernel over blocks x (16 x 16)-sample iterations: blocks x G_PARAMS
// partial gradients (every block writes its whole partial)
void launch_guide_train_mfma(const __bf16 *packed, const float *bias, const float *samples, const uint32_t *n_dev,
                             uint32_t cap, const float *wsum_dev, float *partial, int blocks, hipStream_t st);
// sample ring: ring[0] = next write slot, ring[1] = valid samples (<= cap); a
// gather appended *added samples from slot ring[0] on; advance the ring and
// sum the weights of the valid samples into *wsum in a fixed order
void launch_guide_ring_commit(uint32_t *ring, const uint32_t *added, uint32_t cap, const float *samples,
                              float *wsum, hipStream_t st);
// the fixed-order sum of the per-block partial gradients fused with one Adam
// step (packed / bias as below), in one launch; the sums and their order are
// launch_guide_reduce's (partial is left as it is)
void launch_guide_reduce_adam(const float *partial, int blocks, float *params, float *m, float *v, float lr, float b1,
                              float b2, float eps, int t, hipStream_t st, __bf16 *packed, float *bias);
// packed / bias (may be NULL): also refresh the BF16 weights and biases that
// launch_guide_pack would produce from the updated parameters
void launch_guide_adam(float *params, float *m, float *v, const float *grad, float lr, float b1, float b2, float eps,
                       int t, hipStream_t st, __bf16 *packed = nullptr, float *bias = nullptr);

}  // namespace agx
