// wr_layout.h -- device memory layout shared by the host builder and the kernels.
//
// HBM layout of a scene (all arrays 128-byte aligned, read-only after upload):
//   nodes   : BVH4 nodes, 128 B each = one L2 cache line.  SoA inside the node
//             (the 4 children's x-slabs, then y, then z, then child refs and
//             counts) so one node is eight 16-byte loads and the four box tests
//             run without shuffles.
//   tris    : triangles in leaf order, 48 B each: v0.xyz + original triangle
//             index, v1.xyz + geometry id, v2.xyz + unused.  A leaf is a
//             contiguous run of 1..4 triangles.
//   ltris   : emitter triangles grouped per light, 48 B each: p0.xyz + area-cdf,
//             p1.xyz + area-pdf, p2.xyz + unused.
#pragma once
#include <cstdint>

namespace wr {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmptyChild = 0xFFFFFFFFu;
constexpr int kMaxLeafTris = 4;
constexpr uint32_t kMaxTriangles = 1u << 28;  // leaf refs hold 28-bit triangle offsets
constexpr int kStackDepth = 64;     // traversal stack entries per ray (worst case the builder allows)
constexpr int kLdsStackDepth = 8;   // entries kept in LDS (4 KB per wave); deeper entries spill to a per-lane global region
constexpr int kSpillDepth = kStackDepth - kLdsStackDepth;
constexpr int kMaxTraversalBlocksPerCU = 32;  // bound on traversal grid size (sizes the spill region)

struct alignas(128) Node4 {
    float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
    uint32_t child[4];  // internal: node index; leaf: kLeafBit | (count-1) << 28 | first leaf-order
                        // triangle (first < 2^28); empty: kEmptyChild
    uint32_t count[4];  // leaf: triangle count (1..4); otherwise 0
};
static_assert(sizeof(Node4) == 128, "Node4 must be one cache line");

// The traversal copy of a Node4: 64 B in four 16-byte rows, so a node visit is
// four loads instead of seven (traversal is bound by the L1's rate of lane
// requests, not by arithmetic).  Each child box is stored as 8-bit planes on a
// per-node, per-axis grid of 256 planes: plane q of axis a stands for the FP32 value
// fmaf(q, scale_a, origin_a), computed exactly so by the kernels, and the
// quantiser (k_quantize_nodes) picks every low plane so that this float is <=
// the child's lo and every high plane so that it is >= its hi: the decoded box
// contains the Node4 box, and the kernels' slab test runs on these floats
// unchanged.  Rows: (origin xyz, scale x), (lo_x, hi_x, lo_y, hi_y),
// (lo_z, hi_z, child 0, child 1), (child 2, child 3, scale y, scale z); byte i
// of a plane word is child i's plane.  An empty child keeps kEmptyChild (its
// planes 255 / 0) and is rejected by its ref.
struct alignas(64) NodeQ {
    float org[3];
    float scale_x;       // grid spacing
    uint32_t q[6];       // lo_x, hi_x, lo_y, hi_y, lo_z, hi_z
    uint32_t child[4];   // as Node4::child
    float scale_y, scale_z;
};
static_assert(sizeof(NodeQ) == 64, "NodeQ is four 16-byte rows");

struct TriRecord {  // 48 B
    float v0[3]; uint32_t orig;
    float v1[3]; uint32_t geom;
    float v2[3]; uint32_t pad;
};
static_assert(sizeof(TriRecord) == 48, "TriRecord is 48 bytes");

struct BvhBuildResult {
    Node4 *nodes = nullptr;
    uint32_t n_nodes = 0;
    TriRecord *tris = nullptr;  // leaf order
    uint32_t n_tris = 0;
    uint32_t max_stack = 0;     // worst-case traversal stack entries
    uint32_t depth = 0;
};

// Binned-SAH BVH2 build, collapsed to BVH4 (host, multithreaded).
// verts [nv*3], tris [nt*3], tri_geom [nt].  Returns false on allocation failure.
bool build_bvh4(const float *verts, uint32_t nv, const uint32_t *tris, uint32_t nt, const uint32_t *tri_geom,
                BvhBuildResult *out);
void free_bvh(BvhBuildResult *r);
// Checks the invariants of a built BVH; returns an empty string or the first violation.
const char *verify_bvh4(const BvhBuildResult &r, const float *verts, const uint32_t *tris, uint32_t nt);

}  // namespace wr
