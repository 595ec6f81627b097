// Binned-SAH binary BVH collapsed into a 4-wide BVH (production acceleration
// mode; the layout is described in agx_device.h).
#include "agx_bvh4.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <stdexcept>

namespace agx {
namespace {

constexpr int BINS = 32;
constexpr float C_TRAV = 1.0f;   // SAH cost of a node visit relative to one primitive test
constexpr int MEDIAN_DEPTH = 40; // below this binary depth splits are object medians
constexpr uint32_t PARALLEL_MIN = 1u << 16;
constexpr uint32_t SPARSE_MAX = 32;  // nodes of at most this many references take the sparse SAH sweep

struct Box {
    float lo[3] = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                   std::numeric_limits<float>::infinity()};
    float hi[3] = {-std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                   -std::numeric_limits<float>::infinity()};
    void grow(const float *l, const float *h) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], l[k]);
            hi[k] = std::max(hi[k], h[k]);
        }
    }
    void grow(const Box &b) { grow(b.lo, b.hi); }
    float area() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0) || !(dy >= 0) || !(dz >= 0)) return 0;
        return 2 * (dx * dy + dy * dz + dz * dx);
    }
};

struct BNode {
    Box box;
    int32_t child[2] = {-1, -1};
    uint32_t start = 0, count = 0;  // leaf range in `order` when child[0] < 0
};

// A reference as the builder moves it: bounds, centroid and input index.  The
// records themselves are partitioned (not an index array into the inputs), so
// every pass over a node's references reads memory in order.
struct Item {
    float lo[3], hi[3], c[3];
    uint32_t id;
};

// The binary tree lives in one array with a fixed slot range per subtree: the
// subtree over references [b, e) rooted at slot s uses slots s .. s + 2 (e - b)
// - 2 (a binary tree over n references has at most 2 n - 1 nodes), its left
// child (references [b, m)) is slot s + 1 and its right child slot
// s + 2 (m - b).  Subtrees built on worker threads therefore write disjoint
// slots of the same array, and nothing is copied afterwards.
struct Builder {
    std::vector<Item> items;  // the references, permuted in place as nodes split
    std::vector<BNode> nodes;
    int max_leaf;

    Builder(const std::vector<PrimRef> &r, int ml) : max_leaf(ml) {
        items.resize(r.size());
        for (size_t i = 0; i < r.size(); ++i) {
            Item &it = items[i];
            for (int k = 0; k < 3; ++k) {
                it.lo[k] = r[i].lo[k];
                it.hi[k] = r[i].hi[k];
                it.c[k] = 0.5f * (r[i].lo[k] + r[i].hi[k]);
            }
            it.id = uint32_t(i);
        }
        nodes.resize(2 * r.size() - 1);
    }

    void make_leaf(int32_t self, uint32_t b, uint32_t e) {
        nodes[size_t(self)].start = b;
        nodes[size_t(self)].count = e - b;
    }
    // The binned SAH of a small node without the bin arrays: per axis the
    // references sorted by bin (stably), and only the boundaries between
    // occupied bins evaluated.  Box unions are exact, so each candidate costs
    // what the 32-bin sweep computes for the same two sets, and the lowest bin
    // of equal cost wins as there: the same split, without the fixed cost of
    // clearing and sweeping 3 x 32 bins at every small node.
    void sparse_sah(uint32_t b, uint32_t e, const Box &cb, float *best_cost, int *best_axis, int *best_bin) const {
        const uint32_t n = e - b;
        int key[SPARSE_MAX];
        uint32_t id[SPARSE_MAX];  // positions in items
        Box suf[SPARSE_MAX + 1];  // suf[j] = union of references j .. n - 1 in bin order
        for (int axis = 0; axis < 3; ++axis) {
            const float ext = cb.hi[axis] - cb.lo[axis];
            if (!(ext > 0)) continue;
            const float scale = float(BINS) / ext;
            for (uint32_t i = 0; i < n; ++i) {
                const uint32_t r = b + i;
                int k = int((items[r].c[axis] - cb.lo[axis]) * scale);
                k = std::min(BINS - 1, std::max(0, k));
                uint32_t j = i;  // insertion: ascending bins, equal bins in reference order
                while (j > 0 && key[j - 1] > k) {
                    key[j] = key[j - 1];
                    id[j] = id[j - 1];
                    --j;
                }
                key[j] = k;
                id[j] = r;
            }
            suf[n] = Box();
            for (uint32_t j = n; j-- > 0;) {
                suf[j] = suf[j + 1];
                suf[j].grow(items[id[j]].lo, items[id[j]].hi);
            }
            Box acc;
            for (uint32_t j = 0; j + 1 < n; ++j) {
                acc.grow(items[id[j]].lo, items[id[j]].hi);
                if (key[j] == key[j + 1]) continue;
                const float cost = acc.area() * float(j + 1) + suf[j + 1].area() * float(n - j - 1);
                if (cost < *best_cost) {
                    *best_cost = cost;
                    *best_axis = axis;
                    *best_bin = key[j];
                }
            }
        }
    }

    // bounds and centroid bounds of references [b, e)
    void bounds(uint32_t b, uint32_t e, Box *box, Box *cb) const {
        *box = Box();
        *cb = Box();
        for (uint32_t i = b; i < e; ++i) {
            box->grow(items[i].lo, items[i].hi);
            cb->grow(items[i].c, items[i].c);
        }
    }

    // Builds the subtree over [b, e) rooted at slot self; box and cb are the
    // references' bounds and centroid bounds (the parent's partition pass
    // gathers them).
    void build(int32_t self, uint32_t b, uint32_t e, int depth, const Box &box, const Box &cb) {
        nodes[size_t(self)].box = box;
        const uint32_t n = e - b;
        if (n <= 1 || (depth >= MEDIAN_DEPTH && n <= uint32_t(max_leaf))) {
            make_leaf(self, b, e);
            return;
        }
        // best SAH split over the three axes
        int best_axis = -1, best_bin = -1;
        float best_cost = std::numeric_limits<float>::infinity();
        const float parent_area = box.area();
        if (depth < MEDIAN_DEPTH && n <= SPARSE_MAX) {
            sparse_sah(b, e, cb, &best_cost, &best_axis, &best_bin);
        } else if (depth < MEDIAN_DEPTH) {
            // the three axes' bins in one pass over the references
            Box bins[3][BINS];
            uint32_t cnt[3][BINS] = {};
            float scale[3];
            bool live[3];
            for (int axis = 0; axis < 3; ++axis) {
                const float ext = cb.hi[axis] - cb.lo[axis];
                live[axis] = ext > 0;
                scale[axis] = live[axis] ? float(BINS) / ext : 0.0f;
            }
            for (uint32_t i = b; i < e; ++i) {
                const Item &it = items[i];
                for (int axis = 0; axis < 3; ++axis) {
                    if (!live[axis]) continue;
                    int k = int((it.c[axis] - cb.lo[axis]) * scale[axis]);
                    k = std::min(BINS - 1, std::max(0, k));
                    bins[axis][k].grow(it.lo, it.hi);
                    ++cnt[axis][k];
                }
            }
            for (int axis = 0; axis < 3; ++axis) {
                if (!live[axis]) continue;
                float right_area[BINS];
                uint32_t right_cnt[BINS];
                Box acc;
                uint32_t c = 0;
                for (int k = BINS - 1; k > 0; --k) {
                    acc.grow(bins[axis][k]);
                    c += cnt[axis][k];
                    right_area[k] = acc.area();
                    right_cnt[k] = c;
                }
                acc = Box();
                c = 0;
                for (int k = 0; k < BINS - 1; ++k) {
                    acc.grow(bins[axis][k]);
                    c += cnt[axis][k];
                    if (c == 0 || right_cnt[k + 1] == 0) continue;
                    const float cost = acc.area() * float(c) + right_area[k + 1] * float(right_cnt[k + 1]);
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = axis;
                        best_bin = k;
                    }
                }
            }
        }
        if (best_axis >= 0 && parent_area > 0) {
            const float split_cost = C_TRAV + best_cost / parent_area;
            if (n <= uint32_t(max_leaf) && float(n) <= split_cost) {
                make_leaf(self, b, e);
                return;
            }
        } else if (n <= uint32_t(max_leaf)) {
            make_leaf(self, b, e);  // coincident centroids, few references
            return;
        }
        uint32_t mid;
        Box lbox, lcb, rbox, rcb;  // the children's bounds and centroid bounds
        if (best_axis >= 0) {
            const float ext = cb.hi[best_axis] - cb.lo[best_axis];
            const float scale = float(BINS) / ext;
            const float lo = cb.lo[best_axis];
            auto left_side = [&](const Item &r) {
                int k = int((r.c[best_axis] - lo) * scale);
                k = std::min(BINS - 1, std::max(0, k));
                return k <= best_bin;
            };
            auto to_left = [&](const Item &r) { lbox.grow(r.lo, r.hi); lcb.grow(r.c, r.c); };
            auto to_right = [&](const Item &r) { rbox.grow(r.lo, r.hi); rcb.grow(r.c, r.c); };
            // two-sided (Hoare) partition: scan from the front past left-side
            // references and from the back past right-side ones, swap the pair
            // that stops both scans; each reference's side is settled once, so
            // the children's bounds are gathered on the way
            uint32_t f = b, l = e;
            for (;;) {
                while (f != l && left_side(items[f])) to_left(items[f++]);
                if (f == l) break;
                --l;
                while (f != l && !left_side(items[l])) to_right(items[l--]);
                if (f == l) {
                    to_right(items[f]);  // the right-side reference that stopped the front scan
                    break;
                }
                std::swap(items[f], items[l]);
                to_left(items[f++]);
                to_right(items[l]);
            }
            mid = f;
        } else {
            // object median on the largest centroid extent (deep trees, coincident centroids)
            int axis = 0;
            for (int k = 1; k < 3; ++k)
                if (cb.hi[k] - cb.lo[k] > cb.hi[axis] - cb.lo[axis]) axis = k;
            mid = b + n / 2;
            std::nth_element(items.begin() + b, items.begin() + mid, items.begin() + e,
                             [&](const Item &x, const Item &y) { return x.c[axis] < y.c[axis]; });
            bounds(b, mid, &lbox, &lcb);
            bounds(mid, e, &rbox, &rcb);
        }
        if (mid == b || mid == e) {
            mid = b + n / 2;
            bounds(b, mid, &lbox, &lcb);
            bounds(mid, e, &rbox, &rcb);
        }
        const int32_t left = self + 1, right = self + 2 * int32_t(mid - b);
        nodes[size_t(self)].child[0] = left;
        nodes[size_t(self)].child[1] = right;
        if (n >= PARALLEL_MIN) {
            // left half on a worker thread, right half here
            auto fut = std::async(std::launch::async,
                                  [this, left, b, mid, depth, &lbox, &lcb]() { build(left, b, mid, depth + 1, lbox, lcb); });
            build(right, mid, e, depth + 1, rbox, rcb);
            fut.get();
        } else {
            build(left, b, mid, depth + 1, lbox, lcb);
            build(right, mid, e, depth + 1, rbox, rcb);
        }
    }
};

uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
float bitsf(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Widens a box by a few ulps so that the rounded slab test of a ray that
// meets a primitive exactly on its bound never culls it.
// A child bound as the traversal decodes it: origin + byte x power-of-two step,
// FP32 multiply then add (no contraction), exactly as b4_children evaluates it.
float q_decode(float o, uint32_t q, float step) {
    const float m = float(q) * step;
    return o + m;
}

// Writes one 64-byte node: per axis the children's union low corner and the
// smallest power-of-two step with which byte 255 reaches the union's high
// corner; every child's bounds as bytes whose decoded value encloses the exact
// bound (q = 0 decodes to the origin, q = 255 to the union's top or beyond, so
// the adjustments stop).
void quantize_node(const Box boxes[4], const uint32_t codes[4], vec4 *n) {
    float o[3], step[3];
    uint32_t ebits = 0;
    for (int a = 0; a < 3; ++a) {
        float lo = std::numeric_limits<float>::infinity(), hi = -lo;
        for (int k = 0; k < 4; ++k) {
            if (codes[k] == BVH4_EMPTY) continue;
            lo = std::min(lo, boxes[k].lo[a]);
            hi = std::max(hi, boxes[k].hi[a]);
        }
        if (!(lo <= hi)) lo = hi = 0.0f;
        int e = -126;
        if (hi > lo) {
            int ex = 0;
            (void)std::frexp(double(hi - lo) / 255.0, &ex);
            e = std::max(-126, ex - 1);
        }
        while (e < 127 && q_decode(lo, 255u, std::ldexp(1.0f, e)) < hi) ++e;
        o[a] = lo;
        step[a] = std::ldexp(1.0f, e);
        ebits |= uint32_t(e + 127) << (8 * a);
    }
    uint32_t qw[6] = {0, 0, 0, 0, 0, 0};  // lo.x hi.x lo.y hi.y lo.z hi.z; byte k = child k
    for (int k = 0; k < 4; ++k) {
        if (codes[k] == BVH4_EMPTY) continue;
        for (int a = 0; a < 3; ++a) {
            const double rl = (double(boxes[k].lo[a]) - double(o[a])) / double(step[a]);
            const double rh = (double(boxes[k].hi[a]) - double(o[a])) / double(step[a]);
            uint32_t ql = uint32_t(std::clamp(std::floor(rl), 0.0, 255.0));
            uint32_t qh = uint32_t(std::clamp(std::ceil(rh), 0.0, 255.0));
            while (ql > 0 && q_decode(o[a], ql, step[a]) > boxes[k].lo[a]) --ql;
            while (qh < 255 && q_decode(o[a], qh, step[a]) < boxes[k].hi[a]) ++qh;
            qw[2 * a] |= ql << (8 * k);
            qw[2 * a + 1] |= qh << (8 * k);
        }
    }
    n[0] = vec4{o[0], o[1], o[2], bitsf(ebits)};
    n[1] = vec4{bitsf(codes[0]), bitsf(codes[1]), bitsf(codes[2]), bitsf(codes[3])};
    n[2] = vec4{bitsf(qw[0]), bitsf(qw[1]), bitsf(qw[2]), bitsf(qw[3])};
    n[3] = vec4{bitsf(qw[4]), bitsf(qw[5]), 0.0f, 0.0f};
}

// decoded child boxes and codes of a node (as the traversal sees them)
void decode_node(const vec4 *n, Box boxes[4], uint32_t codes[4]) {
    const uint32_t eb = fbits(n[0].w);
    const float o[3] = {n[0].x, n[0].y, n[0].z};
    const uint32_t qw[6] = {fbits(n[2].x), fbits(n[2].y), fbits(n[2].z), fbits(n[2].w), fbits(n[3].x), fbits(n[3].y)};
    codes[0] = fbits(n[1].x); codes[1] = fbits(n[1].y); codes[2] = fbits(n[1].z); codes[3] = fbits(n[1].w);
    for (int k = 0; k < 4; ++k)
        for (int a = 0; a < 3; ++a) {
            const float step = std::ldexp(1.0f, int((eb >> (8 * a)) & 0xffu) - 127);
            boxes[k].lo[a] = q_decode(o[a], (qw[2 * a] >> (8 * k)) & 0xffu, step);
            boxes[k].hi[a] = q_decode(o[a], (qw[2 * a + 1] >> (8 * k)) & 0xffu, step);
        }
}

void pad(Box *b) {
    for (int k = 0; k < 3; ++k) {
        const float m = std::max(std::fabs(b->lo[k]), std::fabs(b->hi[k]));
        const float e = m * 0x1p-20f + 0x1p-60f;
        b->lo[k] -= e;
        b->hi[k] += e;
    }
}

struct Collapser {
    const std::vector<BNode> &bn;
    Bvh4 *out;
    int max_depth = 0;

    uint32_t leaf_code(const BNode &n) const {
        return BVH4_LEAF | ((n.count - 1) << 27) | n.start;
    }

    // Emits the 4-wide node for binary node `bi` (an interior node or a lone
    // leaf root); returns its index.
    uint32_t emit(int32_t bi, int depth) {
        max_depth = std::max(max_depth, depth);
        int32_t kids[4];
        int nk = 0;
        const BNode &self = bn[size_t(bi)];
        if (self.child[0] < 0) {
            kids[nk++] = bi;
        } else {
            kids[nk++] = self.child[0];
            kids[nk++] = self.child[1];
            while (nk < 4) {
                // open the interior child of largest area (the first of equal areas)
                int pick = -1;
                float best = -1;
                for (int k = 0; k < nk; ++k) {
                    const BNode &c = bn[size_t(kids[k])];
                    if (c.child[0] < 0) continue;
                    const float a = c.box.area();
                    if (a > best) { best = a; pick = k; }
                }
                if (pick < 0) break;
                const BNode &c = bn[size_t(kids[pick])];
                for (int k = nk; k > pick + 1; --k) kids[k] = kids[k - 1];
                kids[pick] = c.child[0];
                kids[pick + 1] = c.child[1];
                ++nk;
            }
        }
        const uint32_t idx = uint32_t(out->nodes.size() / BVH4_NODE_VEC4);
        out->nodes.resize(out->nodes.size() + BVH4_NODE_VEC4, vec4{0, 0, 0, 0});
        uint32_t codes[4] = {BVH4_EMPTY, BVH4_EMPTY, BVH4_EMPTY, BVH4_EMPTY};
        Box boxes[4];
        for (int k = 0; k < nk; ++k) {
            const BNode &c = bn[size_t(kids[k])];
            boxes[k] = c.box;
            pad(&boxes[k]);
            if (c.child[0] < 0) {
                codes[k] = leaf_code(c);
                out->leaves++;
            } else {
                codes[k] = emit(kids[k], depth + 1);
            }
        }
        quantize_node(boxes, codes, &out->nodes[size_t(idx) * BVH4_NODE_VEC4]);
        return idx;
    }
};

}  // namespace

void build_bvh4(const std::vector<PrimRef> &refs, int max_leaf, Bvh4 *out) {
    if (refs.empty()) throw std::invalid_argument("build_bvh4: no primitives");
    if (refs.size() > size_t(BVH4_FIRST_MASK)) throw std::invalid_argument("build_bvh4: too many primitives");
    max_leaf = std::min(BVH4_MAX_LEAF, std::max(1, max_leaf));
    Builder b(refs, max_leaf);
    Box box, cb;
    b.bounds(0, uint32_t(refs.size()), &box, &cb);
    b.build(0, 0, uint32_t(refs.size()), 0, box, cb);
    *out = Bvh4{};
    out->nodes.reserve(b.nodes.size() / 4 * BVH4_NODE_VEC4 + BVH4_NODE_VEC4);
    Collapser c{b.nodes, out};
    c.emit(0, 1);
    out->depth = c.max_depth;
    out->order.resize(b.items.size());
    for (size_t i = 0; i < b.items.size(); ++i) out->order[i] = b.items[i].id;
    // SAH cost of the 4-wide tree relative to the root area
    const float root_area = std::max(b.nodes[0].box.area(), 1e-30f);
    double cost = 0;
    const size_t nn = out->nodes.size() / BVH4_NODE_VEC4;
    for (size_t i = 0; i < nn; ++i) {
        Box boxes[4];
        uint32_t codes[4];
        decode_node(&out->nodes[i * BVH4_NODE_VEC4], boxes, codes);
        for (int k = 0; k < 4; ++k) {
            const uint32_t code = codes[k];
            if (code == BVH4_EMPTY) continue;
            const double a = double(boxes[k].area()) / root_area;
            cost += (code & BVH4_LEAF) ? a * double(((code >> 27) & 15u) + 1) : a * C_TRAV;
        }
    }
    out->sah = cost + C_TRAV;
}

}  // namespace agx
