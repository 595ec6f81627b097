/*
 * atrc_gpu.h — C ABI of the MI355X (gfx950) path tracer that stands in for
 * Atrc's CPU `pt` renderer.
 *
 * Reference interface replaced (paths relative to the Atrc source tree):
 *   - Renderer::render(FilmFilterApplier, Scene&, RendererInteractor&) -> RenderTarget
 *       src/tracer/include/agz/tracer/core/renderer.h:16-64
 *   - PathTracingRenderer / create_pt_renderer / PTRendererParams
 *       src/tracer/src/core/renderer/pt.cpp:8-48, src/tracer/include/agz/tracer/create/renderer.h:10-24
 *   - Scene::closest_intersection / has_intersection (ray-query parity hook)
 *       src/tracer/include/agz/tracer/core/scene.h:42-98, src/tracer/src/core/scene/scene_default.cpp:131-147
 *
 * Atrc's Scene is query-only, so the scene crosses this boundary as plain
 * descriptors (agx_scene_desc) that the host factory fills when it builds the
 * scene from Atrc JSON.  Every pointer is host memory owned by the caller and is
 * only read during the call that receives it.  No call throws; each returns an
 * AGX_* status and agx_last_error() gives the message of the calling thread's
 * last failure.  One context is not re-entrant; distinct contexts are
 * independent.
 */
#ifndef ATRC_GPU_H
#define ATRC_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AGX_ABI_VERSION 13

/* ---- status codes ---------------------------------------------------------------- */
enum {
    AGX_OK              = 0,
    AGX_E_INVALID_ARG   = 1, /* malformed descriptor / argument (ObjectConstructionException) */
    AGX_E_HIP           = 2, /* HIP runtime failure */
    AGX_E_NO_SCENE      = 3, /* render / trace before agx_upload_scene */
    AGX_E_UNSUPPORTED   = 4, /* a feature the GPU path does not implement */
    AGX_E_CANCELLED     = 5, /* *cancel_flag became non-zero; outputs hold the partial image */
    AGX_E_OUT_OF_MEMORY = 6,
    AGX_E_LIMIT         = 7  /* a built structure exceeds a kernel limit (e.g. BVH depth) */
};

/* ---- scene descriptors -------------------------------------------------------------- */

/* Atrc geometry "type" values.  src/factory/src/creator/geometry_creators.cpp */
enum {
    AGX_GEOM_TRIANGLE_BVH = 0, /* "triangle_bvh" / "triangle_bvh_noembree" (world-space triangles) */
    AGX_GEOM_QUAD         = 1, /* "quad": A,B,C,D already in world space */
    AGX_GEOM_TRIANGLE     = 2, /* "triangle": local space + transform */
    AGX_GEOM_SPHERE       = 3, /* "sphere": local space + transform */
    AGX_GEOM_DISK         = 4  /* "disk": local space + transform */
};

/* Atrc material "type" values.  src/factory/src/creator/material_creators.cpp */
enum {
    AGX_MAT_IDEAL_DIFFUSE = 0, /* color = albedo */
    AGX_MAT_METAL         = 1, /* color, eta, k (JSON names), roughness, anisotropic */
    AGX_MAT_GLASS         = 2, /* color = reflection colour, color2 = refraction colour (before sqrt), ior */
    AGX_MAT_MIRROR        = 3, /* color = color_map, eta, k */
    AGX_MAT_IDEAL_BLACK   = 4,
    AGX_MAT_INVISIBLE_SURFACE = 5, /* medium boundary: passes every ray straight through */
    AGX_MAT_DISNEY        = 6  /* color = base_color, color2 = specular_scale, roughness,
                                  anisotropic, ior, and the disney_* fields */
};

enum { AGX_ENV_NONE = 0, AGX_ENV_NATIVE_SKY = 1, AGX_ENV_IBL = 2 };
enum { AGX_TEX2D_CONSTANT = 0, AGX_TEX2D_IMAGE = 1, AGX_TEX2D_CHECKER = 2 };
enum { AGX_FILTER_BOX = 0, AGX_FILTER_GAUSSIAN = 1 };
enum { AGX_AGGREGATE_NATIVE = 0, AGX_AGGREGATE_BVH = 1 };
/* Acceleration structure built at upload.
 *   AGX_ACCEL_ATRC: Atrc's own BVHs and traversal order (triangle_bvh.cpp,
 *     entity_bvh.cpp / native_aggregate.cpp): identical hits, node visits and
 *     primitive tests to Atrc's CPU traversal.
 *   AGX_ACCEL_BVH4: one 4-wide SAH BVH over all primitives, near-first
 *     traversal: identical hits (same primitive, t, u, v) except which of two
 *     primitives at exactly equal t wins; far fewer node visits.
 *   AGX_ACCEL_BVH4_WATERTIGHT: the same BVH4 with the watertight ray/triangle
 *     test of Woop, Benthin and Wald (JCGT 2(1), 2013) on mesh triangles, over
 *     the meshes' own vertex coordinates: a ray through an edge or vertex
 *     shared by two mesh triangles always hits one of them.  Away from edges
 *     the hits are Atrc's up to rounding of t, u, v; analytic entities keep
 *     Atrc's tests. */
enum { AGX_ACCEL_ATRC = 0, AGX_ACCEL_BVH4 = 1, AGX_ACCEL_BVH4_WATERTIGHT = 2 };
enum { AGX_MEDIUM_VOID = 0, AGX_MEDIUM_HOMOGENEOUS = 1, AGX_MEDIUM_HETEROGENEOUS = 2 };
enum { AGX_TEX3D_CONSTANT = 0, AGX_TEX3D_GRID = 1 };
enum { AGX_WRAP_CLAMP = 0, AGX_WRAP_REPEAT = 1, AGX_WRAP_MIRROR = 2 };

/* One mesh vertex; three consecutive vertices form a triangle. */
typedef struct agx_vertex {
    float pos[3];
    float nor[3];
    float uv[2];
} agx_vertex;

/* One Atrc "geometric" entity (src/factory/src/creator/entity_creators.cpp:29-60). */
typedef struct agx_entity_desc {
    int32_t geometry;          /* AGX_GEOM_* */
    int32_t material;          /* index into agx_scene_desc.materials */
    int32_t medium_in;         /* index into media, -1 = void */
    int32_t medium_out;
    float   emit_radiance[3];  /* non-black => the entity is a diffuse area light */
    float   power;             /* user-specified light power, <= 0 = derived */
    int32_t no_denoise;
    uint32_t tri_count;        /* AGX_GEOM_TRIANGLE_BVH */
    const agx_vertex *tris;    /* AGX_GEOM_TRIANGLE_BVH: 3*tri_count vertices in world space */
    float   p[4][3];           /* quad A..D (world) / triangle A..C (local) */
    float   t[4][2];           /* corner texture coordinates */
    float   radius;            /* sphere, disk */
    float   local_to_world[12];/* row-major 3x4 (triangle, sphere, disk) */
    float   world_to_local[12];/* its inverse, row-major 3x4 */
    /* "double_sided" (double_sided.cpp:22-100): a hit from the back turns the
     * geometry's frames around, so the surface shades and emits on both sides;
     * as a light it samples either side with probability 1/2 (pdf halved). */
    int32_t double_sided;
} agx_entity_desc;

/* Atrc Texture3D: "constant3d" or "image3d" (src/tracer/include/agz/tracer/core/texture3d.h,
 * src/tracer/src/core/texture3d/image3d.cpp).  A lookup at uvw applies
 * `transform` (uvw -> texture coordinates), the wrap of each axis, then nearest
 * or trilinear filtering over texel centres, then ^inv_gamma. */
typedef struct agx_texture3d_desc {
    int32_t kind;              /* AGX_TEX3D_* */
    int32_t linear;            /* grid: 1 = trilinear, 0 = nearest */
    int32_t width, height, depth;
    int32_t channels;          /* grid: 1 = real, 3 = spectrum */
    const float *data;         /* grid: depth*height*width*channels floats, x fastest */
    float   value[3];          /* constant value */
    float   transform[12];     /* row-major 3x4 */
    int32_t wrap[3];           /* AGX_WRAP_* for u, v, w */
    float   inv_gamma;
} agx_texture3d_desc;

/* Atrc media (src/factory/src/creator/medium_creators.cpp, src/tracer/src/core/medium/). */
typedef struct agx_medium_desc {
    int32_t kind;              /* AGX_MEDIUM_* */
    int32_t max_scattering_count;
    float   sigma_a[3];        /* homogeneous */
    float   sigma_s[3];
    float   g;                 /* homogeneous Henyey-Greenstein g */
    float   local_to_world[12];/* heterogeneous: unit cube -> world, row-major 3x4 */
    float   world_to_local[12];
    agx_texture3d_desc density;/* heterogeneous: real */
    agx_texture3d_desc albedo; /* heterogeneous: spectrum */
    agx_texture3d_desc g3;     /* heterogeneous: real */
    int32_t white_for_indirect;
} agx_medium_desc;

/* Constant-texture material parameters (image textures are not part of this path). */
typedef struct agx_material_desc {
    int32_t kind;              /* AGX_MAT_* */
    float   color[3];
    float   color2[3];
    float   eta[3];            /* as written in the JSON ("eta") */
    float   k[3];              /* as written in the JSON ("k") */
    float   roughness;
    float   anisotropic;
    float   ior;
    int32_t glass_separate_refraction; /* 1 if color_refraction_map was given separately */
    /* disney (material_creators.cpp:56-112) */
    float   metallic;
    float   transmission;
    float   transmission_roughness;
    float   specular_tint;
    float   sheen;
    float   sheen_tint;
    float   clearcoat;
    float   clearcoat_gloss;
} agx_material_desc;

/* Atrc Texture2D (src/tracer/include/agz/tracer/core/texture2d.h): "constant",
 * "hdr" (RGB float texels) and "checker_board".  A lookup at uv applies
 * `transform` (Texture2DCommonParams::full_transform, inv_u / inv_v / swap_uv
 * folded in), the wrap of each axis, then the texture, then ^inv_gamma. */
typedef struct agx_texture2d_desc {
    int32_t kind;              /* AGX_TEX2D_* */
    int32_t linear;            /* image: 1 = bilinear over texel centres, 0 = nearest */
    int32_t width, height;     /* image */
    const float *data;         /* image: height*width*3 floats, x fastest, row 0 at v = 0 */
    float   value[3];          /* constant texel; checker color1 */
    float   value2[3];         /* checker color2 */
    int32_t grid_count;        /* checker */
    float   transform[6];      /* row-major 2x3 */
    int32_t wrap[2];           /* AGX_WRAP_* for u, v */
    float   inv_gamma;
} agx_texture2d_desc;

typedef struct agx_env_desc {
    int32_t kind;              /* AGX_ENV_* */
    float   top[3];            /* native_sky */
    float   bottom[3];
    float   power;             /* <= 0 = derived */
    agx_texture2d_desc tex;    /* ibl (envir_light/ibl.cpp) */
    int32_t no_importance_sampling; /* ibl: 1 = uniform sphere directions */
} agx_env_desc;

/* Thin-lens camera ("thin_lens"; "pinhole" maps onto it).  src/tracer/src/core/camera/thin_lens.cpp */
typedef struct agx_camera_desc {
    float pos[3];
    float dst[3];
    float up[3];
    float fov_rad;
    float lens_radius;
    float focal_distance;
    float film_aspect;         /* width / height */
} agx_camera_desc;

typedef struct agx_scene_desc {
    uint32_t abi_version;      /* AGX_ABI_VERSION */
    uint32_t entity_count;
    const agx_entity_desc *entities;
    uint32_t material_count;
    const agx_material_desc *materials;
    agx_env_desc env;
    int32_t aggregate;         /* AGX_AGGREGATE_* */
    int32_t max_leaf_size;     /* entity BVH leaf size (default 5) */
    uint32_t medium_count;
    const agx_medium_desc *media; /* referenced by agx_entity_desc.medium_in / medium_out */
    int32_t accel;             /* AGX_ACCEL_* */
} agx_scene_desc;

/* PTRendererParams + the rendering settings that reach the renderer. */
typedef struct agx_render_params {
    int32_t width, height;
    int32_t spp;
    int32_t min_depth;         /* default 5 */
    int32_t max_depth;         /* default 10 */
    int32_t specular_depth;    /* default 20 */
    float   cont_prob;         /* default 0.9 */
    int32_t use_mis;           /* 1: trace_std (NEE with MIS); 0: trace_nomis (no light sampling;
                                  emission and the environment added wherever a path meets them;
                                  path_tracing.cpp:227-381); default 1 */
    float   eps;               /* global EPS, default 3e-4 */
    float   filter_radius;     /* film filter radius as given (default 0.5) */
    uint64_t seed;             /* per-(pixel, sample) PCG32 streams derive from it */
    agx_camera_desc camera;
    int32_t collect_stats;     /* 1: count node visits / triangle tests (slower) */
    int32_t sample_begin;      /* first sample index rendered (progressive rendering), normally 0 */
    /* Path guiding (no counterpart in Atrc; 0 keeps Atrc's estimator). At a
     * Lambert or GGX-metal vertex the continuation direction is drawn
     * from a one-sample mixture: with probability guide_alpha from a learned
     * density (8 von Mises-Fisher lobes predicted by a 4 x 64 MLP from the
     * vertex position, normal and view direction), else from the BSDF; the
     * estimate stays unbiased.  The network lives in the context and keeps
     * learning across renders. */
    int32_t guide;             /* 1: guided continuation sampling */
    float   guide_alpha;       /* in (0, 1); default 0.5 */
    int32_t guide_train;       /* 1: train after every pass of this render (passes of one sample per pixel) */
    int32_t guide_batch;       /* size of the training sample ring: every Adam step trains on the
                                  latest guide_batch samples; at most 1 << 20 (default 65536) */
    int32_t guide_steps;       /* Adam steps per pass (default 4) */
    float   guide_lr;          /* Adam learning rate (default 1e-2) */
    int32_t guide_reset;       /* 1: re-initialise the network from guide_seed first */
    uint32_t guide_seed;
    /* Film filter (film_filter_creators.cpp): AGX_FILTER_BOX (radius - eps; a
     * radius up to 0.5 keeps every sample in its own pixel) or
     * AGX_FILTER_GAUSSIAN (radius, alpha: weight (e^{-alpha x^2} - e^{-alpha r^2})
     * (e^{-alpha y^2} - e^{-alpha r^2})).  Wider filters splat every sample
     * onto all pixels within the radius, and pixels up to the radius outside
     * the image are sampled as well (Atrc's sample_bound_of, framebuffer.h). */
    int32_t filter;            /* AGX_FILTER_*, default box */
    float   filter_alpha;      /* Gaussian alpha */
    /* Atrc's render tile size (PTRendererParams::task_grid_size, default 32;
     * 0 = 32).  Atrc's wide filters splat a sample only within the film grid
     * of the tile that drew it, starting from the tile's first pixel in reach
     * and passing the filter a distance that grows by one per pixel
     * (apply_image_filter, framebuffer.h:49-80); the GPU film does the same,
     * so wide-filter images depend on the tile size as Atrc's do. */
    int32_t task_grid_size;
    /* Path guiding in media: medium scattering points at path depth <= this
     * take the mixture of the learned density and the phase function; deeper
     * ones sample the phase function alone (its weight f / pdf is exactly 1,
     * while the mixture's varies, and the variance of a product of many such
     * weights grows with the scattering count).  0 or negative (default) = no
     * guiding at medium points: in optically thick media, where the light
     * reaching a scattering point is nearly isotropic, the mixture only adds
     * variance and medium records pull the shared network away from the
     * surface distribution. */
    int32_t guide_medium_depth;
} agx_render_params;

/* Counters and timers of the last agx_render / agx_trace_rays call. */
typedef struct agx_stats_t {
    uint64_t closest_rays;     /* closest-hit queries (continuation + MIS-BSDF rays) */
    uint64_t shadow_rays;      /* any-hit queries (NEE shadow rays) */
    uint64_t paths;            /* camera samples traced */
    uint64_t closest_node_visits;  /* BVH node AABB tests in closest-hit queries (collect_stats only) */
    uint64_t closest_tri_tests;    /* triangle tests in closest-hit queries (collect_stats only) */
    uint64_t closest_hits;         /* closest-hit queries that hit (collect_stats only) */
    uint64_t shadow_node_visits;   /* same for any-hit queries */
    uint64_t shadow_tri_tests;
    uint64_t closest_bvh_bytes;    /* algorithmic bytes = visits*node + tests*prim + hits*info */
    uint64_t shadow_bvh_bytes;     /* visits*node + tests*prim */
    double   total_ms;         /* device time of the whole render (HIP events) */
    double   trace_closest_ms; /* device time in the closest-hit kernel */
    double   trace_shadow_ms;  /* device time in the any-hit kernel */
    double   shade_ms;         /* device time in raygen + shade + film kernels (+ guiding features) */
    uint64_t trace_closest_launches;
    uint64_t trace_shadow_launches;
    double   guide_infer_ms;   /* device time of the guiding network's inference kernel */
    double   guide_train_ms;   /* device time of gathering samples + training steps */
    uint64_t guide_samples;    /* training samples gathered from the render's paths */
    uint64_t guide_steps;      /* Adam steps taken by the context's network so far */
    uint64_t guide_rows;       /* vertices the network evaluated */
    uint64_t guide_trained;    /* samples passed through training (ring size x Adam steps) */
    uint32_t node_bytes, prim_bytes, info_bytes; /* record sizes used for the byte counts */
    uint32_t reserved;
    double   host_ms;          /* wall time of the agx_render call */
    double   readback_ms;      /* wall time from the end of the device work to the return
                                  (film resolve, copy to the host outputs) */
} agx_stats_t;

typedef struct agx_ctx agx_ctx;

/* One scene of a batch (agx_render_batch). */
typedef struct agx_batch_item {
    const agx_scene_desc *scene;
    const agx_render_params *params;  /* guide must be 0 */
    float   *rgb;              /* W*H*3 floats as agx_render writes them, or NULL */
    int32_t status;            /* out: AGX_* of this item */
    int32_t worker;            /* out: worker context that rendered it */
    uint64_t rays;             /* out: closest-hit + any-hit queries */
    double  build_ms;          /* out: host build + upload wall time */
    double  render_ms;         /* out: render wall time */
} agx_batch_item;

typedef struct agx_batch_report {
    int32_t workers;           /* scenes in flight at once (agx_render_batch), parallel graph branches
                                  (agx_render_batch_graph) */
    int32_t tuned;             /* 1: chosen by the autotuner */
    double  wall_ms;           /* whole batch, builds included */
    uint64_t rays;
    double  tune_rays_per_s[4];/* autotuner probe throughput: agx_render_batch with 1, 2, 3, 4 workers;
                                  agx_render_batch_graph with 2, 4, 8 branches, then 256-thread
                                  traversal blocks (0 = not tried) */
    /* agx_render_batch_graph only */
    int32_t pass_paths_log2;   /* wavefront pass size of each scene: at most 2^this paths */
    int32_t trace_block;       /* threads per block of the BVH4 traversal launches */
    int32_t groups;            /* graphs launched (one per group of up to 8 scenes) */
    int32_t build_threads;     /* host threads building BVHs */
    double  build_wall_ms;     /* call start to the last scene built */
    double  render_ms;         /* first graph launch to the last image copied out */
    int32_t tune_configs;      /* configurations the autotuner timed */
    int32_t reserved;
} agx_batch_report;

/* Version of this library's ABI (AGX_ABI_VERSION). */
int  agx_abi_version(void);

/* Open a context on a HIP device: its own stream and device buffers. */
int  agx_create(int device_ordinal, agx_ctx **out);

/* Build the BVHs (Atrc's midpoint/median triangle BVH and entity BVH) on the host
 * and copy scene, BVHs and light tables to device memory.  Replaces any
 * previously uploaded scene.  (DefaultSceneCreator + Aggregate::build +
 * Scene::start_rendering, src/factory/src/creator/scene_creators.cpp:15-66) */
int  agx_upload_scene(agx_ctx *ctx, const agx_scene_desc *scene);

/* Render with Atrc `trace_std` semantics (NEE over every light with balance-
 * heuristic MIS, Russian roulette, specular depth extension) and a box film.
 * Outputs are host buffers, row-major, row 0 = film y in [0,1) (Atrc's
 * RenderTarget, before save_to_img flips it): rgb W*H*3, albedo W*H*3,
 * normal W*H*3, denoise W*H; any may be NULL.  cancel_flag (may be NULL) is
 * polled between bounces; a cancelled render returns AGX_E_CANCELLED with the
 * samples finished so far normalised into the outputs. */
int  agx_render(agx_ctx *ctx, const agx_render_params *params,
                const volatile int32_t *cancel_flag,
                float *rgb, float *albedo, float *normal, float *denoise);

/* Ray-query parity hook: n rays in host memory (o, d: n*3 floats; tmin, tmax: n).
 * any_hit = 0: closest hit; outputs entity id (-1 = miss), primitive id (mesh:
 * original triangle index; quad: 0 = ABC, 1 = ACD; others 0), t, and the
 * barycentric (u, v) of the hit triangle.  any_hit = 1: entity_id = 1 if
 * occluded else -1.  Output pointers other than entity_id may be NULL. */
int  agx_trace_rays(agx_ctx *ctx, const float *o, const float *d,
                    const float *tmin, const float *tmax, uint32_t n, int32_t any_hit,
                    int32_t *entity_id, int32_t *prim_id, float *t, float *u, float *v);

/* agx_trace_rays without any device allocation (no counterpart in Atrc): for
 * callers that cannot allocate device memory, such as serving loops and
 * harnesses that forbid hipMalloc.  Same arguments, results and statistics as
 * agx_trace_rays; rays and hits are staged in a fixed device array that the
 * library holds on every GPU it runs on until the process ends
 * (AGX_NOALLOC_MAX_RAYS rays, 84 bytes each, about 11 MB), so n above
 * AGX_NOALLOC_MAX_RAYS returns AGX_E_LIMIT.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (from several threads or
 * contexts) race, and must be serialised by the caller. */
#define AGX_NOALLOC_MAX_RAYS 131072
int  agx_trace_rays_noalloc(agx_ctx *ctx, const float *o, const float *d,
                            const float *tmin, const float *tmax, uint32_t n, int32_t any_hit,
                            int32_t *entity_id, int32_t *prim_id, float *t, float *u, float *v);

int  agx_get_stats(agx_ctx *ctx, agx_stats_t *out);

/* Batched rendering (no counterpart in Atrc, which renders one scene per
 * session): renders every item as agx_upload_scene + agx_render would, with
 * `workers` scenes in flight at once, each on a worker context with its own
 * HIP stream and buffers.  A worker builds its next scene's BVHs on the host
 * while the others keep the GPU busy, and the kernels of different scenes
 * overlap on the device (filling the tail of each wavefront).  workers = 0
 * lets the autotuner choose 1, 2 or 4 from a short probe (the cheapest items
 * at 2 spp).  Images are those agx_render gives for each item alone.  Returns
 * AGX_OK when every item succeeded, else the first failure's status (each
 * item's own status is in item.status). */
int  agx_render_batch(agx_ctx *ctx, agx_batch_item *items, uint32_t n, int32_t workers,
                      agx_batch_report *report);

/* Batched rendering in hipGraph launches (no counterpart in Atrc).  Items are
 * taken smallest build first (triangles) in groups of 8.  Host
 * threads (one per core, at most 16) build the scenes in that order
 * and upload each into one of three sets of 8 pooled contexts, once the group
 * that last used the set has finished; as soon as a group is built, its
 * renders are captured into one hipGraph with `branches` parallel branches
 * (the costliest scene first onto the least loaded branch; a branch renders
 * its scenes one after another) and launched, so later scenes build while
 * earlier groups render.  A group's images go to pinned host memory on the
 * stream right after its graph and to the items' buffers once it is done.
 * A captured render launches every bounce up to the depth bound (max_depth +
 * 2, plus specular_depth in scenes with delta surfaces; kernels of an empty
 * queue exit at once) and counts its rays on the device.  branches = 0: an
 * autotuner times the first group at 2 spp for 2, 4 and 8 branches, passes
 * of 2^21 / 2^22 / 2^23 paths and 128- / 256-thread traversal blocks
 * (coordinate descent; later groups build meanwhile) and renders the batch
 * with the fastest; otherwise `branches` branches, passes of at most 2^22
 * paths and 128-thread blocks.  Images are bit-identical to agx_render's for
 * each item alone.  item.worker = the item's branch; item.render_ms = its
 * group's graph time.  After a successful call the pooled contexts (their
 * path buffers and last scenes) stay with ctx for its next batch, so a call
 * neither starts by allocating them again nor ends by freeing them;
 * agx_destroy(ctx) releases them. */
int  agx_render_batch_graph(agx_ctx *ctx, agx_batch_item *items, uint32_t n, int32_t branches,
                            agx_batch_report *report);

/* Message of the calling thread's last failed call ("" if none). */
const char *agx_last_error(void);

void agx_destroy(agx_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* ATRC_GPU_H */
