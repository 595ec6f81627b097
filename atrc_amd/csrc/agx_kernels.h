// Wavefront buffers and kernel launch entry points of the gfx950 path tracer.
#pragma once

#include <hip/hip_runtime.h>

#include "agx_device.h"

namespace agx {

// Per-path state (structure of arrays, N = paths in flight).
struct PathBuffers {
    float4 *coef;        // [N] path throughput
    float4 *radiance;    // [N] accumulated radiance
    float4 *gb_albedo;   // [N] (albedo.xyz, denoise)
    float4 *gb_normal;   // [N]
    unsigned long long *rng;     // [N] PCG32 state (the increment is derived from the path id)
    unsigned long long *rng_tr;  // [N] second stream: ratio-tracking transmittance draws
    int4 *depth;         // [N] (depth, s_depth, medium scattering count, 0)
    uint32_t *flags;     // [N] PF_* bits
    float4 *mis_a;       // [N] MIS-BSDF record: (coef.xyz, |cos|)
    float4 *mis_b;       // [N]                  (f.xyz, pdf)
    float4 *nee;         // [N*L] (coef * contribution.xyz, valid)
    float4 *ray_o;       // [2N] (o.xyz, t_min); 2p = continuation, 2p+1 = MIS-BSDF ray
    float4 *ray_d;       // [2N] (d.xyz, t_max)
    float4 *hit;         // [2N] (t, u, v, entity bits), entity = -1 on miss
    int32_t *hit_prim;   // [2N] primitive slot (mesh: global BVH slot; quad: 0/1)
    float4 *sh_o;        // [N*L] shadow rays (o.xyz, t_max)
    float4 *sh_d;        // [N*L] (d.xyz, t_min)
    uint8_t *occluded;   // [N*L]
    uint32_t *q_closest; // [2N] ray indices to trace (closest hit)
    uint32_t *q_shadow;  // [N*L] shadow ray indices
    uint32_t *q_in;      // [N] paths to shade this bounce
    uint32_t *q_out;     // [N] paths to shade next bounce
    const uint32_t *cnt_in;  // [4] sizes of the queues consumed this bounce: closest, shadow, paths
    uint32_t *cnt_out;       // [4] sizes of the queues produced for the next bounce
    float *film_sum;     // [10*WH] rgb, albedo, normal, denoise
    float *film_comp;    // [10*WH] Neumaier compensation
    float *film_count;   // [WH]
    // path guiding (RenderConst::guide)
    float *g_feat;       // [N*8] network inputs of the vertices in q_in order
    float *g_out;        // [N*40] decoded mixture of those vertices (8 x (mu.xyz, kappa), 8 weights)
    float *g_rec;        // [guide_rec_paths(N) * GUIDE_MAX_REC * GUIDE_REC_FLOATS] training records
    int2 *g_rec_state;   // [guide_rec_paths(N)] (records made, record of the previous vertex or -1)
    float4 *g_ms;        // [3N] guided renders in media: the medium sample of the vertex about to be
                         // shaded, taken before inference: (pos.xyz, g), (throughput.xyz, happened), (wr.xyz, 0)
};

// Path guiding: one path in 2^guide_rec_shift keeps training records of its
// first GUIDE_MAX_REC guided vertices: inputs[8], direction[3], mixture pdf,
// throughput after the vertex[3], incident radiance estimate[3], luminance of
// f |cos| of the sampled direction.  The stride keeps at most GUIDE_REC_PATHS
// recording paths per pass (GUIDE_REC_PATHS * GUIDE_MAX_REC * 76 B = 160 MB):
// light that reaches a vertex only through a narrow opening is found by few
// paths, so the record budget is large.
constexpr int GUIDE_REC_PATHS = 1 << 18;
constexpr int GUIDE_MAX_REC = 8;
constexpr int GUIDE_REC_FLOATS = 19;
constexpr int GUIDE_ROW = 40;
// smallest shift that leaves at most GUIDE_REC_PATHS recording paths among n
inline uint32_t guide_rec_shift(size_t n) {
    uint32_t s = 0;
    while ((n >> s) > size_t(GUIDE_REC_PATHS)) ++s;
    return s;
}
// record slots needed by a pass of n paths at that shift
inline size_t guide_rec_paths(size_t n) { return n ? ((n - 1) >> guide_rec_shift(n)) + 1 : 0; }

enum : uint32_t {
    PF_CONT = 1u,        // continuation ray 2p is queued
    PF_MIS = 2u,         // MIS-BSDF ray 2p+1 is queued
    PF_PENDING = 4u,     // NEE / MIS contributions await resolution
    PF_KEEP = 8u,        // the camera sample lands inside the box filter
    PF_MIS_DELTA = 16u,  // the MIS-BSDF sample was a delta lobe
    PF_FRESH = 32u,      // camera path not shaded yet: radiance 0, depth (1, 1, 0, 0) and the
                         // depth-1 throughput are implied (k_raygen does not store them)
    PF_MS = 64u,         // k_medium_pass sampled this vertex's medium segment into g_ms
    PF_MIS_MED_SHIFT = 16u,  // bits 16..31: 1 + medium index the MIS-BSDF ray travels through (0 = void)
};
// key of the transmittance stream (initstate = splitmix(seed ^ key ^ splitmix(seq)))
constexpr unsigned long long TR_STREAM_KEY = 0xA5A5F00DC0FFEE11ull;

struct RenderConst {
    int32_t W, H;
    int32_t min_depth, max_depth, specular_depth;
    float cont_prob;
    float eps;
    float filter_r;      // box radius - eps
    unsigned long long seed;
    int64_t work_base;   // first work item of this pass; work = layer * W*H + pixel
    int32_t n_paths;     // paths in this pass
    int32_t sample_begin;
    int32_t L;           // number of lights
    int32_t guide;       // 1: guided continuation sampling
    int32_t guide_media; // 1: guided render of a scene with media (k_medium_pass runs before inference)
    int32_t guide_medium_depth;  // medium scattering points at path depth <= this are guided
    float guide_alpha;   // probability of drawing from the learned density
    int32_t guide_record;  // 1: guided vertices of the recording paths keep training records
    uint32_t guide_rec_shift;  // recording paths: (p + guide_rec_rot) % 2^shift == 0, record slot p >> shift
    uint32_t guide_rec_rot;    // varies per pass so that every pixel records in turn
    int32_t pad2;
    // wide film filters (box wider than a pixel, Gaussian): every sample splats
    // onto the pixels of its render tile (tile x tile pixels, Atrc's
    // task_grid_size) as Atrc's apply_image_filter reaches them; the sampled
    // grid extends border0 pixels before and border1 pixels after the image on
    // each axis (EW x EH pixels)
    int32_t splat;
    int32_t tile;
    int32_t nomis;         // 1: trace_nomis (use_mis = 0): no light sampling, emission at every hit
    int32_t filter_kind;   // AGX_FILTER_*
    int32_t border0, border1;
    int32_t EW, EH;
    float filter_alpha, filter_exp;  // Gaussian: alpha and e^{-alpha r^2}
    float4 cam_pos, cam_x, cam_y, cam_z;  // camera frame
    float film_w, film_h, lens_r, focal;
};

void launch_raygen(const DevScene &s, const RenderConst &rc, const PathBuffers &pb, hipStream_t st);
// The launch grids cover the capacity (max_n); each kernel reads its real
// queue size from pb.cnt_in, so the host never waits between bounces.
void launch_trace_closest(const DevScene &s, const PathBuffers &pb, uint32_t max_n, bool stats, DevCounters *cnt,
                          hipStream_t st);
void launch_trace_shadow(const DevScene &s, const PathBuffers &pb, uint32_t max_n, bool stats, DevCounters *cnt,
                         hipStream_t st);
void launch_shade(const DevScene &s, const RenderConst &rc, const PathBuffers &pb, uint32_t max_n, hipStream_t st);
// queues of the next bounce from the paths k_shade shaded (at most max_n of them);
// blk holds queue_compact_blocks(max_n) entries
uint32_t queue_compact_blocks(uint32_t max_n);
// *acc += cin[0] + cin[1] (rays of the coming bounce), on the stream
void launch_count_rays(const uint32_t *cin, unsigned long long *acc, hipStream_t st);
void launch_queue_compact(const PathBuffers &pb, uint32_t max_n, int L, uint4 *blk, hipStream_t st);
// network inputs of the vertices about to be shaded (q_in order)
void launch_guide_features(const DevScene &s, const RenderConst &rc, const PathBuffers &pb, uint32_t max_n,
                           hipStream_t st);
// guided renders in media: samples the medium segment of every vertex about to
// be shaded (so that a scattering point can get its own network evaluation);
// k_shade then uses the stored sample
void launch_medium_pass(const DevScene &s, const RenderConst &rc, const PathBuffers &pb, uint32_t max_n,
                        hipStream_t st);
// training samples (inputs, direction, weight) from the records of finished paths;
// appends at most cap samples, counts them in *count and sums their weights in *wsum
void launch_guide_gather(const RenderConst &rc, const PathBuffers &pb, float *samples, uint32_t cap,
                         uint32_t *count, const uint32_t *ring, uint32_t *blk, hipStream_t st);
// blocks of the sample gather (blk of launch_guide_gather holds this many entries)
constexpr uint32_t GUIDE_GATHER_BLOCKS = uint32_t(GUIDE_REC_PATHS) * GUIDE_MAX_REC / 256 + 1;
void launch_reset_counters(uint32_t *c, hipStream_t st);
void launch_copy_calibration(const float4 *src, float4 *dst, uint64_t n, hipStream_t st);
// rays of the closest-hit (which = 0) or any-hit (1) queue appended to o/d (at most cap)
void launch_capture_rays(const PathBuffers &pb, int which, float4 *o, float4 *d, uint32_t *count, uint32_t cap,
                         uint32_t max_n, hipStream_t st);
void launch_film(const RenderConst &rc, const PathBuffers &pb, hipStream_t st);
void launch_film_clear(const PathBuffers &pb, int64_t wh, hipStream_t st);
void launch_film_resolve(const PathBuffers &pb, int64_t wh, float *rgb, float *albedo, float *normal, float *denoise,
                         hipStream_t st);
// parity hook: rays in device memory, results to device memory
// Staging of agx_trace_rays_noalloc: a file-scope device array of
// NOALLOC_MAX_RAYS rays (84 B each: packed rays, raw inputs, outputs) that the
// library never allocates or frees.  noalloc_scratch_init resolves its address
// on the current device (once per device; call with that device current);
// noalloc_scratch returns it, or nullptr before init.
constexpr uint32_t NOALLOC_MAX_RAYS = 1u << 17;
constexpr size_t NOALLOC_BYTES_PER_RAY = 84;
hipError_t noalloc_scratch_init(int device);
void *noalloc_scratch(int device);
// trace_rays over raw host-layout inputs staged in the scratch: packs
// (o, tmin) / (d, tmax), clears cnt with a kernel, then traces
void launch_trace_rays_staged(const DevScene &s, void *scratch, uint32_t n, int any_hit, DevCounters *cnt,
                              hipStream_t st);
void launch_trace_rays(const DevScene &s, const float4 *o, const float4 *d, uint32_t n, int any_hit, int32_t *ent,
                       int32_t *prim, float *t, float *u, float *v, DevCounters *cnt, hipStream_t st);

}  // namespace agx
