// Device-resident scene and wavefront state of the gfx950 path tracer.
//
// HBM layout (all arrays 16-byte aligned, structure-of-arrays where a kernel
// streams one field over many paths):
//   entities   DevEntity[n]            one record per Atrc entity (288 B)
//   materials  DevMaterial[m]          constant-texture BSDF parameters, pre-derived
//   blas_nodes float4[2*N]             Atrc triangle-BVH nodes: (low.xyz, start) (high.xyz, end|right)
//   blas_prims float4[3*P]             (a.xyz, original id) (b-a, 0) (c-a, 0)        48 B / triangle
//   blas_info  float4[6*P]             interpolated normal / uv deltas and frame      96 B / triangle
//   tlas_nodes float4[2*T]             entity BVH: (low.xyz, left|leaf-start) (high.xyz, right|leaf-end)
//   tlas_prims int32[]                 entity indices referenced by TLAS leaves
//   alias_prob/alias_idx               per-mesh triangle-area alias tables (emissive meshes)
//   lights     int32[L]                Atrc light order: -1 = environment, else entity index
#pragma once

#include <stdint.h>

namespace agx {

// 16-byte POD shared by host code and kernels (loads as one dwordx4)
struct alignas(16) vec4 {
    float x, y, z, w;
};

constexpr uint32_t INTERIOR = 0xffffffffu;   // BLAS node: start == INTERIOR marks an interior node
constexpr uint32_t TLAS_LEAF = 0x80000000u;  // TLAS node: left word has this bit for leaves
constexpr int MAX_LIGHTS = 16;
// BVH4 (production acceleration mode): a node is 4 vec4 = 64 B:
//   [0] (origin.xyz, step exponents: byte a = e_a + 127, step_a = 2^e_a)
//   [1] child codes (uint bits)
//   [2] (lo.x, hi.x, lo.y, hi.y), [3] (lo.z, hi.z, 0, 0): each a uint with byte k
//       = bound of child k; bound = origin + float(byte) * step (FP32 multiply,
//       then add), always enclosing the child's exact box.
// Child code: BVH4_EMPTY, an interior node index (bit 31 clear), or a leaf
// BVH4_LEAF | (count - 1) << 27 | first primitive.  A primitive is 3 vec4:
//   (A, entity), (B - A, slot), (C - A, tie rank); slot = index into
//   blas_prims/blas_info of a mesh triangle, or BVH4_ANALYTIC for a whole
//   analytic entity; tie rank = DevEntity::tie_rank of the entity.  In the
//   watertight mode a mesh triangle is (A, entity), (B, slot), (C, tie rank)
//   with the mesh's own vertex coordinates, and a quad is two primitives
//   (A, entity), (B, BVH4_QUAD_ABC), (C, rank) and (A, entity), (C, BVH4_QUAD_ACD), (D, rank).
constexpr int BVH4_NODE_VEC4 = 4;
constexpr uint32_t BVH4_EMPTY = 0xffffffffu;
constexpr uint32_t BVH4_LEAF = 0x80000000u;
constexpr uint32_t BVH4_FIRST_MASK = 0x07ffffffu;
constexpr int BVH4_MAX_LEAF = 8;
constexpr uint32_t BVH4_ANALYTIC = 0xffffffffu;
// watertight mode: a quad enters the BVH4 as its two triangles over the
// descriptor's own corners, (A, B, C) and (A, C, D)
constexpr uint32_t BVH4_QUAD_ABC = 0xfffffffeu;
constexpr uint32_t BVH4_QUAD_ACD = 0xfffffffdu;
constexpr int BVH4_STACK_SMALL = 64;   // traversal stack entries; the large variant
constexpr int BVH4_STACK_LARGE = 192;  // serves trees deeper than (64 - 1) / 3 levels
// The newest entries of a lane's BVH4 stack live in an LDS ring, older ones in
// private memory: 8 (code, entry distance) pairs of the closest-hit traversal
// (16 KiB per 256-thread block, so eight waves per SIMD still fit the CU's
// 160 KiB) and 16 codes of the any-hit traversal.
constexpr int BVH4_RING_CLOSEST = 8;
constexpr int BVH4_RING_ANY = 16;
constexpr int BLAS_STACK = 96;               // >= max BLAS depth + 2 (checked at upload)
constexpr int TLAS_STACK = 64;
constexpr int LDS_STACK = 32;                // LDS stack column per thread (TLAS + BLAS entries)

// DevEntity::flags
constexpr int32_t EF_NO_DENOISE = 1;
constexpr int32_t EF_DOUBLE_SIDED = 2;  // double_sided.cpp: back-side hits turn the frames around

struct DevEntity {
    int32_t kind;          // AGX_GEOM_*
    int32_t material;
    int32_t light;         // -1 or index into the light list
    int32_t flags;         // EF_*
    int32_t med_in, med_out;  // media on the -z / +z side of the geometric normal, -1 = void
    int32_t tie_rank;      // order in which Atrc's aggregate visits the entity: on an exactly equal
                           // hit distance the later-visited entity wins (native: index; BVH: leaf order)
    int32_t pad_m1;
    vec4 a;                // quad/triangle: a.xyz ; sphere/disk: (radius, radius^2, world_radius, ratio)
    vec4 b_a, c_a, d_a;    // edges (quad: world space; triangle: local space)
    vec4 x_abc, x_acd, z;  // geometric frame vectors
    vec4 t_ab;             // (t_a.x, t_a.y, t_b_a.x, t_b_a.y)
    vec4 t_cd;             // (t_c_a.x, t_c_a.y, t_d_a.x, t_d_a.y)
    vec4 l2w[3];           // rows of local_to_world
    vec4 w2l[3];           // rows of world_to_local
    vec4 emit;             // (radiance.xyz, surface area)
    uint32_t node_base, prim_base, prim_count, alias_base;  // mesh ranges in the global arrays
    vec4 sample_prob;      // (quad: P(ABC), 0, 0, 0)
};
static_assert(sizeof(DevEntity) == 320, "DevEntity layout");

// Texture3D (constant or grid); grids live in DevScene::grids
struct DevTex3D {
    int32_t kind, linear, w, h;
    int32_t d, ch;
    uint32_t offset;       // first texel of the grid in DevScene::grids (or grids16)
    float inv_gamma;
    int32_t wrap[3];
    int32_t f16;           // 1: texels stored as FP16 in DevScene::grids16 (every texel exactly representable)
    vec4 value;
    vec4 xf[3];            // rows of the uvw transform
};

// Texture2D (constant / hdr image / checker board) of an ibl environment
struct DevTex2D {
    int32_t kind, linear, w, h;
    int32_t wrap[2];
    int32_t grid_count;
    float inv_gamma;
    vec4 value, value2;
    vec4 xf[2];            // rows of the uv transform (2x3)
};

struct DevMedium {
    int32_t kind;          // AGX_MEDIUM_*
    int32_t max_scat;
    int32_t white_for_indirect;
    int32_t pad;
    vec4 sa, ss, st;       // homogeneous coefficients (xyz)
    vec4 misc;             // (g, max density, 1 / max density, 0)
    vec4 w2l[3];           // heterogeneous: world -> unit cube
    DevTex3D density, albedo, g3;
};

struct DevMaterial {
    int32_t kind;          // AGX_MAT_*
    int32_t pad0, pad1, pad2;
    vec4 albedo;           // BSDF::albedo()
    vec4 coef;             // Lambert albedo / pi
    vec4 color;            // metal colour
    vec4 eta2;             // conductor (eta_in / eta_out)^2 ; mirror uses m_eta2
    vec4 k2;               // conductor (k / eta_out)^2
    vec4 misc;             // (ax, ay, ior, 0); disney (ax, ay, ior, roughness)
    vec4 c_refl, c_refr;   // glass; disney (tint, specular_scale)
    vec4 d0;               // disney (metallic, specular_tint, sheen, sheen_tint)
    vec4 d1;               // disney (clearcoat, transmission, trans_ax, trans_ay)
    vec4 d2;               // disney (clearcoat_roughness, w_diffuse, w_transmission, w_specular)
    vec4 d3;               // disney (w_clearcoat, 0, 0, 0)
};

struct DevScene {
    const DevEntity *ents;
    const DevMaterial *mats;
    const vec4 *blas_nodes;
    const vec4 *blas_prims;
    const vec4 *blas_info;
    const vec4 *tlas_nodes;
    const int32_t *tlas_prims;
    const float *alias_prob;
    const int32_t *alias_idx;
    const DevMedium *media;
    const float *grids;
    const uint16_t *grids16;  // FP16 texels of the grids whose values FP16 holds exactly
    int32_t n_ent;
    int32_t aggregate;     // AGX_AGGREGATE_*
    int32_t n_lights;
    int32_t env_kind;
    int32_t lds_stack;     // 1: TLAS depth + BLAS depth fit the LDS stack column
    int32_t accel;         // AGX_ACCEL_*
    int32_t bvh4_large_stack;  // 1: the BVH4 is too deep for BVH4_STACK_SMALL
    uint32_t features;     // SF_* the scene uses (media, Disney, delta materials)
    int32_t trace_block;   // threads per block of the BVH4 traversal kernels (128 or 256; host-side choice)
    const vec4 *b4_nodes;
    const vec4 *b4_prims;
    int32_t lights[MAX_LIGHTS];
    vec4 sky_top, sky_bottom;
    // ibl environment (AGX_ENV_IBL): texture, and the importance sampler over
    // env_pw x env_ph patches (normalised patch probabilities + alias table)
    DevTex2D env_tex;
    const float *env_texels;
    const float *env_probs;
    const float *env_alias_p;
    const int32_t *env_alias_i;
    int32_t env_pw, env_ph;
    vec4 world;            // (centre.xyz, radius)
};

// Features a shading kernel is compiled with; scenes without a feature use the
// variant without its code (fewer registers, higher occupancy).
enum : uint32_t {
    SF_MEDIA = 1u,    // participating media
    SF_DISNEY = 2u,   // Disney BSDF
    SF_DELTA = 4u,    // glass, mirror, invisible_surface
    SF_GUIDE = 8u,    // path guiding
    SF_IBL = 16u,     // image-based environment light
    SF_ALL = 31u,
};

// per-launch counters (collect_stats)
struct DevCounters {
    unsigned long long node_visits;
    unsigned long long tri_tests;
    unsigned long long hits;
    unsigned long long pad;
};

}  // namespace agx
