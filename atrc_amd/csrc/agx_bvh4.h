// Production acceleration structure: one 4-wide binned-SAH BVH over every
// primitive of the scene (each mesh triangle, each analytic entity).
//
// Atrc builds a midpoint/median binary BVH per mesh and a median-split entity
// BVH on top (triangle_bvh.cpp:122-223, entity_bvh.cpp:46-116); that pair is
// kept for the parity mode.  Atrc has no instancing (triangle_bvh meshes are
// transformed to world space at construction, triangle_bvh.cpp:533-558), so a
// single-level BVH over world-space primitives answers the same closest-hit and
// any-hit queries with far fewer node visits: SAH splits, 4 children per node
// (one 128-byte fetch tests four boxes), near-first traversal with entry-distance
// culling.  Primitive tests use Atrc's own intersection arithmetic, so a hit
// reports the same primitive, t, u, v; only which of two primitives at exactly
// the same t wins can differ from Atrc's traversal order.
#pragma once

#include <cstdint>
#include <vector>

#include "agx_device.h"

namespace agx {

// World-space bound of one primitive reference.
struct PrimRef {
    float lo[3], hi[3];
};

struct Bvh4 {
    std::vector<vec4> nodes;      // BVH4_NODE_VEC4 vec4 per node, root at 0
    std::vector<uint32_t> order;  // leaf order: references to PrimRef indices
    int depth = 0;                // node levels (root = 1)
    uint64_t leaves = 0;
    double sah = 0;               // SAH cost (traversal 1, intersection 1) relative to the root area
};

// Builds the BVH4 over refs (at least one).  max_leaf in [1, BVH4_MAX_LEAF].
void build_bvh4(const std::vector<PrimRef> &refs, int max_leaf, Bvh4 *out);

}  // namespace agx
