// Host-side construction of the device scene from agx_scene_desc.
#pragma once

#include <string>
#include <vector>

#include "../../include/atrc_gpu.h"
#include "agx_device.h"

namespace agx {

struct HostScene {
    std::vector<DevEntity> ents;
    std::vector<DevMaterial> mats;
    std::vector<vec4> blas_nodes, blas_prims, blas_info, tlas_nodes;
    std::vector<int32_t> tlas_prims, alias_idx;
    std::vector<float> alias_prob;
    std::vector<int32_t> lights;
    std::vector<DevMedium> media;
    std::vector<float> grids;
    std::vector<uint16_t> grids16;
    int32_t accel = AGX_ACCEL_ATRC;
    std::vector<vec4> b4_nodes, b4_prims;  // AGX_ACCEL_BVH4
    int b4_depth = 0;
    uint64_t b4_leaves = 0;
    double b4_sah = 0;
    int32_t aggregate = AGX_AGGREGATE_NATIVE;
    int32_t env_kind = AGX_ENV_NONE;
    vec4 sky_top{0, 0, 0, 0}, sky_bottom{0, 0, 0, 0}, world{0, 0, 0, 1};
    // ibl environment: texture, its texels, and the patch sampler
    DevTex2D env_tex{};
    std::vector<float> env_texels, env_probs, env_alias_p;
    std::vector<int32_t> env_alias_i;
    int32_t env_pw = 0, env_ph = 0;
    int max_blas_depth = 0;
    int max_tlas_depth = 0;
    uint64_t total_tris = 0;
};

// Validates the descriptor and builds every structure.  Returns an AGX_* status;
// *err receives the message on failure.
int build_host_scene(const agx_scene_desc &desc, HostScene *out, std::string *err);

// Atrc triangle BVH of one mesh, flattened: nodes (2 vec4 each, local indices),
// prims (3 vec4, w of the first = original triangle id), info (6 vec4).
// Returns the maximum node depth.
int build_triangle_bvh(const agx_vertex *tris, uint32_t n, std::vector<vec4> *nodes,
                       std::vector<vec4> *prims, std::vector<vec4> *info, double *surface_area);

}  // namespace agx
