/*
 * atrc_gpu_debug.h — host-only inspection hooks of libatrc_gpu.so, used by the
 * CPU test-suite to check the scene build without a GPU.  They have no
 * counterpart in Atrc; they expose what agx_upload_scene builds from the same
 * descriptor (the BVHs of src/tracer/src/core/geometry/triangle_bvh.cpp:122-304
 * and src/tracer/src/core/aggregate/entity_bvh.cpp:46-116).
 */
#ifndef ATRC_GPU_DEBUG_H
#define ATRC_GPU_DEBUG_H

#include "atrc_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct agx_debug_summary {
    uint32_t entities, lights;
    uint32_t blas_nodes, blas_prims, tlas_nodes;
    uint32_t max_blas_depth, max_tlas_depth;
    uint64_t device_bytes;
    float world_centre_x, world_centre_y, world_centre_z, world_radius;
    /* AGX_ACCEL_BVH4 scenes: node count, node levels, leaves, SAH cost */
    uint64_t bvh4_nodes, bvh4_leaves;
    uint32_t bvh4_depth, pad_;
    double bvh4_sah;
    /* 3D texture texels stored as FP32, and as FP16 (grids whose every texel
     * FP16 holds exactly) */
    uint64_t grid_texels_f32, grid_texels_f16;
} agx_debug_summary;

/* Atrc triangle BVH of one mesh.  nodes_out (may be NULL): node_count * 8
 * floats, per node (low.xyz, start bits, high.xyz, end-or-right bits);
 * prims_out (may be NULL): n * 12 floats (a.xyz, original id bits, b-a, 0, c-a, 0). */
int agx_debug_triangle_bvh(const agx_vertex *tris, uint32_t n, uint32_t *node_count, uint32_t *max_depth,
                           float *nodes_out, float *prims_out);

/* Build (without uploading) the device scene and report its sizes. */
int agx_debug_scene_summary(const agx_scene_desc *desc, agx_debug_summary *out);

/* The BVH4 of an AGX_ACCEL_BVH4 descriptor (layout in atrc_amd/csrc/agx_device.h):
 * *n_nodes nodes of 32 floats and *n_prims primitives of 12 floats.  Call with
 * NULL outputs to get the counts, then with buffers of that size. */
int agx_debug_bvh4(const agx_scene_desc *desc, uint64_t *n_nodes, uint64_t *n_prims, float *nodes_out,
                   float *prims_out);

/* Path-guiding network (no counterpart in Atrc), exposed for its numerics tests.
 * params: AGX_GUIDE_PARAMS floats (per layer l = 0..4: W_l [in][out] row-major,
 * then b_l; layer sizes 256-64-64-64-64-32).
 * infer: inputs n x 8 in [0, 1] -> out n x 40: 8 x (mu.xyz, kappa), 8 weights. */
#define AGX_GUIDE_PARAMS 31008
int agx_debug_guide_infer(agx_ctx *ctx, const float *params, const float *inputs, uint32_t n, float *out);
/* One training step on n samples of 12 floats (8 inputs, direction xyz, weight):
 * grad_out (may be NULL) receives the gradient of sum_s w_s wscale (-log q(dir_s));
 * params, m, v are updated in place by one Adam step (t >= 1) unless lr == 0. */
int agx_debug_guide_train(agx_ctx *ctx, float *params, float *m, float *v, const float *samples, uint32_t n,
                          float wscale, float lr, int t, float *grad_out);

/* The same gradient from the matrix-core training path that renders use (BF16
 * operands, FP32 accumulation); no parameter update. */
int agx_debug_guide_grad_mfma(agx_ctx *ctx, const float *params, const float *samples, uint32_t n, float wscale,
                              float *grad_out);

/* Run one plain device-to-device copy kernel (k_copy_calibration) of `bytes`
 * bytes on the context's stream and wait for it: a known-size reference for
 * calibrating hardware traffic counters.  Returns its device time in *ms. */
int agx_debug_copy_kernel(agx_ctx *ctx, uint64_t bytes, double *ms);

/* Ray replay: renders `params` once (as agx_render, image discarded) while
 * capturing every closest-hit and any-hit query the wavefront issues, in queue
 * order (at most max_rays of each), then traces the captured sets again with
 * the traversal kernel alone (no shading, no counters): the BVH-bound ray
 * throughput of the same ray distribution.  out[0..1] = rays captured
 * (closest, any-hit), out_ms[0..1] = average device time of one replay of
 * each set (1 warm-up, then `reps` timed), out_ms[2..3] = the render's own
 * closest-hit / any-hit kernel time.  AGX_E_LIMIT when a set overflows. */
int agx_debug_replay_rays(agx_ctx *ctx, const agx_render_params *params, uint32_t max_rays, int reps,
                          uint64_t out[2], double out_ms[4]);

#ifdef __cplusplus
}
#endif

#endif /* ATRC_GPU_DEBUG_H */
