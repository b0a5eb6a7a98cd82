// wr_bvh_gpu.hip -- BVH build on the GPU: PLOC clustering (Meister and Bittner,
// "Parallel Locally-Ordered Clustering for Bounding Volume Hierarchy
// Construction", IEEE TVCG 24(12), 2018) over Morton-ordered triangle boxes
// gives a binary tree whose leaves are chosen by the SAH; it is then collapsed
// top-down, one tree level per launch pair, into the 128-byte Node4 layout of
// wr_layout.h that the traversal kernels read.  Which binary nodes become the
// children of each 4-wide node is chosen by the SAH too (Ylitie, Karras and
// Laine, "Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide
// BVHs", HPG 2017, section 4.1), by a table that each merge fills from its two
// children's, the same rule as the host builder's collapse.
//
// Replaces rtcCommitScene's BVH8/Triangle4 SAH build (reference:
// src/accel/embreeaccel.h:466-469; Embree kernels/common/scene.cpp:120-140).
// Any valid BVH gives the same closest hit, so the tree shape only sets the
// traversal cost.
//
// The build is deterministic: the radix sort is stable and every allocation is
// a prefix sum.  The nearest-neighbour search ranks the pairs (i, j) by the key
// (merged area, -hash(min(i,j), max(i,j), iteration), min, max), a strict total
// order that is the same from both ends of a pair, so the smallest pair is
// always mutual (at least one merge per iteration); the hash turns runs of
// equal areas (duplicate triangles) into random matchings of logarithmic depth
// instead of a chain.
//
// Device memory of the binary tree: 2n-1 nodes of 60 B (box lo.xyz + SAH cost,
// box hi.xyz + triangle count with the leaf mark, two child ids, the collapse
// table: four costs and one word of decisions); nodes 0..n-1 are the triangles
// in Morton order, inner nodes follow in merge order, so a child always has a
// smaller id than its parent.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <cfloat>
#include <chrono>
#include <cstring>

#include "wr_host.h"

namespace wr {
namespace {

constexpr int kBB = 256;              // threads per block of the build kernels
constexpr int kPlocRadius = 16;       // default nearest-neighbour search window on each side
constexpr int kPlocRadiusMax = 64;
constexpr float kCt = 1.0f, kCi = 1.0f;  // SAH costs of a box test and a triangle test (as the host builder)
constexpr uint32_t kLeafMark = 0x80000000u;  // in a node's count word: the subtree becomes one Node4 leaf
constexpr int kBoundsBlocks = 1024;
constexpr uint32_t kNone = 0xFFFFFFFFu;   // no node / unused child slot

// Collapse table of binary node x: wc[x] = (c1, c2, c3, c4), ci = least SAH cost
// of covering x's subtree with at most i children of a 4-wide node; wd[x] =
// decisions: bit 0 = with one child, x is one wide leaf (else one wide inner
// node); bits 2(i-1) .. 2(i-1)+1 for i = 2..4 = s: 0 = as with i-1 children,
// else s children from the left subtree and i-s from the right; bits 8..9 =
// the split of x's own four slots when x is a wide inner node.
constexpr uint32_t kWideLeaf = 1u;
__device__ __forceinline__ int wide_split(uint32_t w, int i) { return int(w >> (2 * (i - 1))) & 3; }
__device__ __forceinline__ int wide_top(uint32_t w) { return int(w >> 8) & 3; }

__device__ __forceinline__ float half_area(float3 lo, float3 hi) {
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    if (dx < 0.f || dy < 0.f || dz < 0.f) return 0.f;
    return dx * dy + dy * dz + dz * dx;
}

__device__ __forceinline__ float3 f3(const float *v, uint32_t i) {
    return make_float3(v[3 * size_t(i)], v[3 * size_t(i) + 1], v[3 * size_t(i) + 2]);
}

__device__ __forceinline__ void tri_box(const float *verts, const uint32_t *tris, uint32_t t, float3 &lo, float3 &hi) {
    const float3 a = f3(verts, tris[3 * size_t(t)]), b = f3(verts, tris[3 * size_t(t) + 1]),
                 c = f3(verts, tris[3 * size_t(t) + 2]);
    lo = make_float3(fminf(fminf(a.x, b.x), c.x), fminf(fminf(a.y, b.y), c.y), fminf(fminf(a.z, b.z), c.z));
    hi = make_float3(fmaxf(fmaxf(a.x, b.x), c.x), fmaxf(fmaxf(a.y, b.y), c.y), fmaxf(fmaxf(a.z, b.z), c.z));
}

// ---------------------------------------------------------------- centroid bounds
// Per-block min/max of the triangle-box centroids -> partial[6 * block].
__global__ __launch_bounds__(kBB) void k_centroid_bounds(const float *verts, const uint32_t *tris, uint32_t n,
                                                         float *partial) {
    __shared__ float red[6][kBB];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * kBB + threadIdx.x; i < n; i += gridDim.x * kBB) {
        float3 lo, hi;
        tri_box(verts, tris, i, lo, hi);
        const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
        for (int k = 0; k < 3; k++) {
            v[k] = fminf(v[k], c[k]);
            v[3 + k] = fmaxf(v[3 + k], c[k]);
        }
    }
    for (int k = 0; k < 6; k++) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = kBB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 3; k++) {
                red[k][threadIdx.x] = fminf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
                red[3 + k][threadIdx.x] = fmaxf(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// One block: reduces `blocks` partials into cbox[6].
__global__ __launch_bounds__(kBB) void k_centroid_final(const float *partial, int blocks, float *cbox) {
    __shared__ float red[6][kBB];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < blocks; b += kBB)
        for (int k = 0; k < 3; k++) {
            v[k] = fminf(v[k], partial[6 * b + k]);
            v[3 + k] = fmaxf(v[3 + k], partial[6 * b + 3 + k]);
        }
    for (int k = 0; k < 6; k++) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = kBB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 3; k++) {
                red[k][threadIdx.x] = fminf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
                red[3 + k][threadIdx.x] = fmaxf(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) cbox[threadIdx.x] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------- Morton codes
// 21 bits -> every third bit of 63.
__device__ __forceinline__ uint64_t spread21(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | x << 32) & 0x001F00000000FFFFull;
    x = (x | x << 16) & 0x001F0000FF0000FFull;
    x = (x | x << 8) & 0x100F00F00F00F00Full;
    x = (x | x << 4) & 0x10C30C30C30C30C3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__global__ __launch_bounds__(kBB) void k_morton(const float *verts, const uint32_t *tris, uint32_t n,
                                                const float *cbox, uint64_t *keys, uint32_t *vals) {
    const uint32_t i = blockIdx.x * kBB + threadIdx.x;
    if (i >= n) return;
    float3 lo, hi;
    tri_box(verts, tris, i, lo, hi);
    const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
        const float ext = cbox[3 + k] - cbox[k];
        const float u = ext > 0.f ? (c[k] - cbox[k]) / ext : 0.f;
        q[k] = uint32_t(fminf(fmaxf(u * 2097152.f, 0.f), 2097151.f));
    }
    keys[i] = spread21(q[0]) | spread21(q[1]) << 1 | spread21(q[2]) << 2;
    vals[i] = i;
}

// Leaves: node j = the j-th triangle in Morton order.
__global__ __launch_bounds__(kBB) void k_leaves(const float *verts, const uint32_t *tris, uint32_t n,
                                                const uint32_t *order, float4 *nlo, float4 *nhi, uint2 *kids,
                                                float4 *wc, uint32_t *wd, uint32_t *clusters) {
    const uint32_t j = blockIdx.x * kBB + threadIdx.x;
    if (j >= n) return;
    const uint32_t t = order[j];
    float3 lo, hi;
    tri_box(verts, tris, t, lo, hi);
    const float c = kCi * half_area(lo, hi);
    nlo[j] = make_float4(lo.x, lo.y, lo.z, c);
    nhi[j] = make_float4(hi.x, hi.y, hi.z, __uint_as_float(1u | kLeafMark));
    kids[j] = make_uint2(kNone, t);
    wc[j] = make_float4(c, c, c, c);
    wd[j] = kWideLeaf;
    clusters[j] = j;
}

// ---------------------------------------------------------------- PLOC iteration
// Symmetric pair priority for ties of the merged area.
__device__ __forceinline__ uint32_t pair_hash(uint32_t lo, uint32_t hi, uint32_t seed) {
    uint64_t x = (uint64_t(hi) << 32 | lo) ^ (uint64_t(seed) * 0x9E3779B97F4A7C15ull);
    x ^= x >> 33;
    x *= 0xFF51AFD7ED558CCDull;
    x ^= x >> 33;
    x *= 0xC4CEB9FE1A85EC53ull;
    x ^= x >> 33;
    return uint32_t(x);
}

// nearest[i]: the cluster in [i - R, i + R] whose union with cluster i has the
// smallest surface area; equal areas go to the higher pair hash, then to the
// lower index.
__global__ __launch_bounds__(kBB) void k_nearest(const uint32_t *clusters, uint32_t m, const float4 *nlo,
                                                 const float4 *nhi, int R, uint32_t seed, uint32_t *nearest) {
    __shared__ float4 slo[kBB + 2 * kPlocRadiusMax], shi[kBB + 2 * kPlocRadiusMax];
    const int64_t base = int64_t(blockIdx.x) * kBB - R;
    for (int s = threadIdx.x; s < kBB + 2 * R; s += kBB) {
        const int64_t g = base + s;
        if (g >= 0 && g < int64_t(m)) {
            const uint32_t id = clusters[g];
            slo[s] = nlo[id];
            shi[s] = nhi[id];
        }
    }
    __syncthreads();
    const int64_t i = int64_t(blockIdx.x) * kBB + threadIdx.x;
    if (i >= int64_t(m)) return;
    const int me = threadIdx.x + R;
    const float3 alo = make_float3(slo[me].x, slo[me].y, slo[me].z), ahi = make_float3(shi[me].x, shi[me].y, shi[me].z);
    float best = INFINITY;
    uint32_t bh_best = 0;
    int64_t bj = -1;
    for (int o = -R; o <= R; o++) {  // ascending j: on a full tie the first (lower) j stays
        const int64_t j = i + o;
        if (o == 0 || j < 0 || j >= int64_t(m)) continue;
        const float4 bl = slo[me + o], bh = shi[me + o];
        const float3 lo = make_float3(fminf(alo.x, bl.x), fminf(alo.y, bl.y), fminf(alo.z, bl.z));
        const float3 hi = make_float3(fmaxf(ahi.x, bh.x), fmaxf(ahi.y, bh.y), fmaxf(ahi.z, bh.z));
        const float a = half_area(lo, hi);
        if (a > best) continue;
        const uint32_t h = o < 0 ? pair_hash(uint32_t(j), uint32_t(i), seed) : pair_hash(uint32_t(i), uint32_t(j), seed);
        if (a < best || h > bh_best) { best = a; bh_best = h; bj = j; }
    }
    // non-finite areas (validated input has none) still pair with a neighbour
    if (bj < 0) bj = i > 0 ? i - 1 : i + 1;
    nearest[i] = uint32_t(bj);
}

// flags[i]: bit 32 = cluster i merges with its mutual nearest neighbour (the
// lower index of the pair owns the merge); bit 0 = cluster i survives.
__global__ __launch_bounds__(kBB) void k_merge_flags(const uint32_t *nearest, uint32_t m, uint64_t *flags) {
    const uint32_t i = blockIdx.x * kBB + threadIdx.x;
    if (i >= m) return;
    const uint32_t j = nearest[i];
    const bool mutual = nearest[j] == i;
    const bool merge = mutual && i < j;
    const bool keep = !(mutual && i > j);
    flags[i] = (uint64_t(merge) << 32) | uint64_t(keep);
}

// Creates the merged nodes (ids node_base + merge rank) with their SAH cost and
// leaf decision, and compacts the surviving clusters in order.
__global__ __launch_bounds__(kBB) void k_merge(const uint32_t *clusters, const uint32_t *nearest, uint32_t m,
                                               const uint64_t *flags, const uint64_t *incl, uint32_t node_base,
                                               float4 *nlo, float4 *nhi, uint2 *kids, float4 *wc, uint32_t *wd,
                                               uint32_t *out) {
    const uint32_t i = blockIdx.x * kBB + threadIdx.x;
    if (i >= m) return;
    const uint64_t f = flags[i];
    if (!(f & 1u)) return;
    const uint64_t s = incl[i];
    const uint32_t pos = uint32_t(s) - 1u;
    if (!(f >> 32)) {
        out[pos] = clusters[i];
        return;
    }
    const uint32_t a = clusters[i], b = clusters[nearest[i]];
    const uint32_t id = node_base + uint32_t(s >> 32) - 1u;
    const float4 la = nlo[a], ha = nhi[a], lb = nlo[b], hb = nhi[b];
    const float3 lo = make_float3(fminf(la.x, lb.x), fminf(la.y, lb.y), fminf(la.z, lb.z));
    const float3 hi = make_float3(fmaxf(ha.x, hb.x), fmaxf(ha.y, hb.y), fmaxf(ha.z, hb.z));
    const uint32_t cnt = (__float_as_uint(ha.w) & ~kLeafMark) + (__float_as_uint(hb.w) & ~kLeafMark);
    const float A = half_area(lo, hi);
    const float inner = kCt * A + la.w + lb.w;
    const float leaf_cost = kCi * float(cnt) * A;
    const bool leaf = cnt <= uint32_t(kMaxLeafTris) && leaf_cost <= inner;
    nlo[id] = make_float4(lo.x, lo.y, lo.z, leaf ? leaf_cost : inner);
    nhi[id] = make_float4(hi.x, hi.y, hi.z, __uint_as_float(cnt | (leaf ? kLeafMark : 0u)));
    kids[id] = make_uint2(a, b);
    // collapse table (a binary SAH leaf stays one wide leaf)
    if (leaf) {
        wc[id] = make_float4(leaf_cost, leaf_cost, leaf_cost, leaf_cost);
        wd[id] = kWideLeaf;
    } else {
        const float4 ca = wc[a], cb = wc[b];
        const float cl[4] = {ca.x, ca.y, ca.z, ca.w}, cr[4] = {cb.x, cb.y, cb.z, cb.w};
        float dist[5];
        uint32_t dk[5];
#pragma unroll
        for (int j = 2; j <= 4; j++) {  // k = 1 first: a valid split even if a cost is not finite
            dist[j] = cl[0] + cr[j - 2];
            dk[j] = 1;
#pragma unroll
            for (int k = 2; k < j; k++) {
                const float v = cl[k - 1] + cr[j - k - 1];
                if (v < dist[j]) { dist[j] = v; dk[j] = uint32_t(k); }
            }
        }
        const float wide_inner = kCt * A + dist[4];
        const float wide_leaf = cnt <= uint32_t(kMaxLeafTris) ? leaf_cost : INFINITY;
        float c[4];
        uint32_t w = wide_leaf <= wide_inner ? kWideLeaf : 0u;
        c[0] = fminf(wide_inner, wide_leaf);
#pragma unroll
        for (int i = 2; i <= 4; i++) {
            const bool take = dist[i] < c[i - 2];
            c[i - 1] = take ? dist[i] : c[i - 2];
            w |= (take ? dk[i] : 0u) << (2 * (i - 1));
        }
        wc[id] = make_float4(c[0], c[1], c[2], c[3]);
        wd[id] = w | dk[4] << 8;
    }
    out[pos] = id;
}

// ---------------------------------------------------------------- collapse to Node4
__device__ __forceinline__ uint32_t count_of(const float4 *nhi, uint32_t x) {
    return __float_as_uint(nhi[x].w) & ~kLeafMark;
}

// Queue entry e = (binary node, first leaf-order triangle of its subtree,
// traversal stack entries that may be live when the Node4 is visited).
// Picks the node's children by its collapse table, left to right.
__global__ __launch_bounds__(kBB) void k_expand(const uint4 *queue, uint32_t qn, const float4 *nlo,
                                                const float4 *nhi, const uint2 *kids, const uint32_t *wd,
                                                uint4 *ekid, uint4 *eoff, uint32_t *einner) {
    const uint32_t e = blockIdx.x * kBB + threadIdx.x;
    if (e >= qn) return;
    const uint4 q = queue[e];
    uint32_t kid[4] = {kNone, kNone, kNone, kNone}, off[4] = {0, 0, 0, 0};
    int k = 0;
    const uint32_t wq = wd[q.x];
    if (wq & kWideLeaf) {  // only the root can be a leaf here
        kid[0] = q.x;
        off[0] = q.y;
        k = 1;
    } else {
        // pending (subtree, slots): the slots of all pending entries and of the
        // current one sum to at most 4, each at least 1, so at most 3 wait
        uint32_t pn[4], pi[4];
        const uint2 c = kids[q.x];
        const uint32_t t = uint32_t(wide_top(wq));
        pn[0] = c.y; pi[0] = 4u - t;
        pn[1] = c.x; pi[1] = t;
        int np = 2;
        uint32_t o = q.y;
        while (np > 0) {
            np--;
            uint32_t x = pn[np], i = pi[np];
            for (;;) {
                const uint32_t s = i > 1u ? uint32_t(wide_split(wd[x], int(i))) : 0u;
                if (i <= 1u) {
                    if (k < 4) { kid[k] = x; off[k] = o; k++; }
                    o += count_of(nhi, x);
                    break;
                }
                if (s == 0u || s >= i) { i--; continue; }
                const uint2 g = kids[x];
                if (np < 4) { pn[np] = g.y; pi[np] = i - s; np++; }
                x = g.x;
                i = s;
            }
        }
    }
    // slot order = decreasing area (stable): any-hit traversal visits the slots
    // in order, and a larger box is likelier to hold an occluder
    float area[4];
    for (int i = 0; i < k; i++) {
        const float4 lo = nlo[kid[i]], hi = nhi[kid[i]];
        area[i] = half_area(make_float3(lo.x, lo.y, lo.z), make_float3(hi.x, hi.y, hi.z));
    }
    for (int i = 1; i < k; i++)
        for (int j = i; j > 0 && area[j] > area[j - 1]; j--) {
            const float ta = area[j]; area[j] = area[j - 1]; area[j - 1] = ta;
            const uint32_t tk = kid[j]; kid[j] = kid[j - 1]; kid[j - 1] = tk;
            const uint32_t to = off[j]; off[j] = off[j - 1]; off[j - 1] = to;
        }
    uint32_t inner = 0;
    for (int i = 0; i < k; i++) inner += (wd[kid[i]] & kWideLeaf) ? 0u : 1u;
    ekid[e] = make_uint4(kid[0], kid[1], kid[2], kid[3]);  // unused slots kNone
    eoff[e] = make_uint4(off[0], off[1], off[2], off[3]);
    einner[e] = inner;
}

struct CollapseArgs {
    const uint4 *queue;
    uint32_t qn, level_base, next_base;
    const uint4 *ekid, *eoff;
    const uint32_t *einner, *eincl;
    const float4 *nlo, *nhi;
    const uint2 *kids;
    const uint32_t *wd;                  // collapse decisions
    uint32_t n;                          // triangles = binary leaves
    const float *verts;
    const uint32_t *tris, *tri_geom;
    Node4 *out;
    float4 *tri_out;                     // leaf-order TriRecords (3 float4 each)
    uint32_t *leaf_orig;                 // leaf order -> input triangle
    uint4 *qnext;
    uint32_t *max_stack;
};

__global__ __launch_bounds__(kBB) void k_collapse(CollapseArgs a) {
    const uint32_t e = blockIdx.x * kBB + threadIdx.x;
    if (e >= a.qn) return;
    const uint4 q = a.queue[e];
    const uint4 kv = a.ekid[e], ov = a.eoff[e];
    const uint32_t kid[4] = {kv.x, kv.y, kv.z, kv.w}, off[4] = {ov.x, ov.y, ov.z, ov.w};
    uint32_t k = 0;
    for (int i = 0; i < 4; i++) k += kid[i] != kNone ? 1u : 0u;
    uint32_t slot = a.eincl[e] - a.einner[e];
    float lo[3][4], hi[3][4];
    uint32_t child[4], count[4];
    for (int i = 0; i < 4; i++) {
        for (int d = 0; d < 3; d++) { lo[d][i] = FLT_MAX; hi[d][i] = -FLT_MAX; }
        child[i] = kEmptyChild;
        count[i] = 0;
    }
    for (uint32_t i = 0; i < k; i++) {
        const uint32_t x = kid[i];
        const float4 bl = a.nlo[x], bh = a.nhi[x];
        lo[0][i] = bl.x; lo[1][i] = bl.y; lo[2][i] = bl.z;
        hi[0][i] = bh.x; hi[1][i] = bh.y; hi[2][i] = bh.z;
        if (a.wd[x] & kWideLeaf) {
            const uint32_t cnt = count_of(a.nhi, x);
            child[i] = kLeafBit | ((cnt - 1u) << 28) | off[i];
            count[i] = cnt;
            // the subtree's triangles, left to right, at off[i] ..; at most kMaxLeafTris leaves
            uint32_t st[kMaxLeafTris + 1];
            int sp = 0;
            st[sp++] = x;
            uint32_t p = off[i];
            while (sp > 0 && p < off[i] + cnt) {
                const uint32_t v = st[--sp];
                if (v < a.n) {
                    const uint32_t t = a.kids[v].y;
                    const uint32_t i0 = a.tris[3 * size_t(t)], i1 = a.tris[3 * size_t(t) + 1],
                                   i2 = a.tris[3 * size_t(t) + 2];
                    float4 *r = a.tri_out + 3 * size_t(p);
                    r[0] = make_float4(a.verts[3 * size_t(i0)], a.verts[3 * size_t(i0) + 1], a.verts[3 * size_t(i0) + 2],
                                       __uint_as_float(t));
                    r[1] = make_float4(a.verts[3 * size_t(i1)], a.verts[3 * size_t(i1) + 1], a.verts[3 * size_t(i1) + 2],
                                       __uint_as_float(a.tri_geom[t]));
                    r[2] = make_float4(a.verts[3 * size_t(i2)], a.verts[3 * size_t(i2) + 1], a.verts[3 * size_t(i2) + 2],
                                       0.f);
                    a.leaf_orig[p] = t;
                    p++;
                } else if (sp + 2 <= kMaxLeafTris + 1) {
                    const uint2 g = a.kids[v];
                    st[sp++] = g.y;
                    st[sp++] = g.x;
                }
            }
        } else {
            child[i] = a.next_base + slot;
            a.qnext[slot] = make_uint4(x, off[i], q.z + k - 1u, 0u);
            slot++;
        }
    }
    atomicMax(a.max_stack, q.z + k);
    float4 *o = reinterpret_cast<float4 *>(a.out + a.level_base + e);
    o[0] = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
    o[1] = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
    o[2] = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
    o[3] = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
    o[4] = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
    o[5] = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
    reinterpret_cast<uint4 *>(o)[6] = make_uint4(child[0], child[1], child[2], child[3]);
    reinterpret_cast<uint4 *>(o)[7] = make_uint4(count[0], count[1], count[2], count[3]);
}

inline unsigned grid_of(uint64_t n) { return unsigned((n + kBB - 1) / kBB); }

// Device allocations of one build, released on every return path once the
// build's stream has finished with them.
struct Scratch {
    hipStream_t stream = nullptr;
    std::vector<void *> ptrs;
    template <typename T>
    hipError_t alloc(T **p, size_t count) {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(count * sizeof(T), 256));
        if (e == hipSuccess) ptrs.push_back(q);
        *p = static_cast<T *>(q);
        return e;
    }
    ~Scratch() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : ptrs) (void)hipFree(p);
    }
};

}  // namespace

int build_bvh4_gpu(const float *d_verts, const uint32_t *d_tris, const uint32_t *d_tri_geom, uint32_t n,
                   hipStream_t stream, int radius, uint32_t *leaf_to_orig, GpuBvhResult *res) {
    *res = GpuBvhResult{};
    const int R = radius > 0 ? std::min(radius, kPlocRadiusMax) : kPlocRadius;
    const auto t0 = std::chrono::steady_clock::now();
    if (n == 0) {
        Node4 nd;
        for (int i = 0; i < 4; i++) {
            nd.lo_x[i] = nd.lo_y[i] = nd.lo_z[i] = FLT_MAX;
            nd.hi_x[i] = nd.hi_y[i] = nd.hi_z[i] = -FLT_MAX;
            nd.child[i] = kEmptyChild;
            nd.count[i] = 0;
        }
        WR_HIP(hipMalloc(&res->nodes, sizeof(Node4)));
        WR_HIP(hipMalloc(&res->tris, 128));
        WR_HIP(hipMemcpy(res->nodes, &nd, sizeof(Node4), hipMemcpyHostToDevice));
        res->n_nodes = 1;
        res->max_stack = 1;
        res->depth = 1;
        return WR_OK;
    }
    Scratch sc;
    sc.stream = stream;
    const size_t nn = 2 * size_t(n) - 1;
    float *partial, *cbox;
    uint64_t *keys, *keys2, *flags, *incl;
    uint32_t *vals, *order, *cl, *cl2, *nearest, *einner, *eincl, *max_stack, *leaf_orig;
    float4 *nlo, *nhi, *wc;
    uint32_t *wd;
    uint2 *kids;
    uint4 *q, *q2, *ekid, *eoff;
    Node4 *nodes_tmp;
    const size_t node_cap = std::max<size_t>(1, n);
    WR_HIP(sc.alloc(&partial, 6 * kBoundsBlocks));
    WR_HIP(sc.alloc(&cbox, 6));
    WR_HIP(sc.alloc(&keys, n));
    WR_HIP(sc.alloc(&keys2, n));
    WR_HIP(sc.alloc(&vals, n));
    WR_HIP(sc.alloc(&order, n));
    WR_HIP(sc.alloc(&nlo, nn));
    WR_HIP(sc.alloc(&nhi, nn));
    WR_HIP(sc.alloc(&kids, nn));
    WR_HIP(sc.alloc(&wc, nn));
    WR_HIP(sc.alloc(&wd, nn));
    WR_HIP(sc.alloc(&cl, n));
    WR_HIP(sc.alloc(&cl2, n));
    WR_HIP(sc.alloc(&nearest, n));
    WR_HIP(sc.alloc(&flags, n));
    WR_HIP(sc.alloc(&incl, n));
    WR_HIP(sc.alloc(&q, node_cap));
    WR_HIP(sc.alloc(&q2, node_cap));
    WR_HIP(sc.alloc(&ekid, node_cap));
    WR_HIP(sc.alloc(&eoff, node_cap));
    WR_HIP(sc.alloc(&einner, node_cap));
    WR_HIP(sc.alloc(&eincl, node_cap));
    WR_HIP(sc.alloc(&max_stack, 1));
    WR_HIP(sc.alloc(&leaf_orig, n));
    WR_HIP(sc.alloc(&nodes_tmp, node_cap));
    // library temporary storage, sized for the largest call
    size_t sort_bytes = 0, scan64_bytes = 0, scan32_bytes = 0;
    WR_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, keys, keys2, vals, order, n, 0, 63, stream));
    WR_HIP(rocprim::inclusive_scan(nullptr, scan64_bytes, flags, incl, size_t(n), rocprim::plus<uint64_t>(), stream));
    WR_HIP(rocprim::inclusive_scan(nullptr, scan32_bytes, einner, eincl, node_cap, rocprim::plus<uint32_t>(), stream));
    void *tmp;
    const size_t tmp_bytes = std::max(sort_bytes, std::max(scan64_bytes, scan32_bytes));
    WR_HIP(sc.alloc(reinterpret_cast<uint8_t **>(&tmp), tmp_bytes));

    // Morton order
    const int nb = int(std::min<uint64_t>(kBoundsBlocks, grid_of(n)));
    k_centroid_bounds<<<nb, kBB, 0, stream>>>(d_verts, d_tris, n, partial);
    k_centroid_final<<<1, kBB, 0, stream>>>(partial, nb, cbox);
    k_morton<<<grid_of(n), kBB, 0, stream>>>(d_verts, d_tris, n, cbox, keys, vals);
    WR_HIP(hipGetLastError());
    size_t bytes = sort_bytes;
    WR_HIP(rocprim::radix_sort_pairs(tmp, bytes, keys, keys2, vals, order, n, 0, 63, stream));
    k_leaves<<<grid_of(n), kBB, 0, stream>>>(d_verts, d_tris, n, order, nlo, nhi, kids, wc, wd, cl);
    WR_HIP(hipGetLastError());

    // PLOC
    uint32_t m = n, node_base = n, iters = 0;
    uint64_t total = 0;
    while (m > 1) {
        k_nearest<<<grid_of(m), kBB, 0, stream>>>(cl, m, nlo, nhi, R, iters, nearest);
        k_merge_flags<<<grid_of(m), kBB, 0, stream>>>(nearest, m, flags);
        WR_HIP(hipGetLastError());
        bytes = scan64_bytes;
        WR_HIP(rocprim::inclusive_scan(tmp, bytes, flags, incl, size_t(m), rocprim::plus<uint64_t>(), stream));
        k_merge<<<grid_of(m), kBB, 0, stream>>>(cl, nearest, m, flags, incl, node_base, nlo, nhi, kids, wc, wd,
                                                cl2);
        WR_HIP(hipGetLastError());
        WR_HIP(hipMemcpyAsync(&total, incl + (m - 1), sizeof total, hipMemcpyDeviceToHost, stream));
        WR_HIP(hipStreamSynchronize(stream));
        const uint32_t merges = uint32_t(total >> 32), kept = uint32_t(total);
        if (merges == 0 || kept != m - merges || node_base + merges > nn)
            return set_error(WR_ERR_DEVICE, "GPU BVH build: clustering made no progress (non-finite vertices?)");
        node_base += merges;
        m = kept;
        std::swap(cl, cl2);
        iters++;
    }
    if (node_base != nn) return set_error(WR_ERR_DEVICE, "GPU BVH build: node count mismatch");
    uint32_t root = 0;
    WR_HIP(hipMemcpyAsync(&root, cl, sizeof root, hipMemcpyDeviceToHost, stream));

    // collapse, level by level (breadth-first node order: children after parents)
    WR_HIP(hipMalloc(&res->tris, std::max<size_t>(sizeof(TriRecord) * n, 128)));
    const uint4 q0 = make_uint4(root, 0u, 0u, 0u);
    WR_HIP(hipStreamSynchronize(stream));
    WR_HIP(hipMemcpyAsync(q, &q0, sizeof q0, hipMemcpyHostToDevice, stream));
    WR_HIP(hipMemsetAsync(max_stack, 0, sizeof(uint32_t), stream));
    uint32_t qn = 1, level_base = 0, depth = 0;
    while (qn > 0) {
        const uint32_t next_base = level_base + qn;
        k_expand<<<grid_of(qn), kBB, 0, stream>>>(q, qn, nlo, nhi, kids, wd, ekid, eoff, einner);
        WR_HIP(hipGetLastError());
        bytes = scan32_bytes;
        WR_HIP(rocprim::inclusive_scan(tmp, bytes, einner, eincl, size_t(qn), rocprim::plus<uint32_t>(), stream));
        uint32_t inner_total = 0;
        CollapseArgs ca{q, qn, level_base, next_base, ekid, eoff, einner, eincl, nlo, nhi, kids, wd, n,
                        d_verts, d_tris, d_tri_geom, nodes_tmp, reinterpret_cast<float4 *>(res->tris), leaf_orig,
                        q2, max_stack};
        // this level's nodes end at next_base (checked by the previous level); its
        // children must fit behind them
        WR_HIP(hipMemcpyAsync(&inner_total, eincl + (qn - 1), sizeof inner_total, hipMemcpyDeviceToHost, stream));
        WR_HIP(hipStreamSynchronize(stream));
        if (uint64_t(next_base) + inner_total > node_cap)
            return set_error(WR_ERR_DEVICE, "GPU BVH build: node capacity exceeded");
        k_collapse<<<grid_of(qn), kBB, 0, stream>>>(ca);
        WR_HIP(hipGetLastError());
        level_base = next_base;
        qn = inner_total;
        std::swap(q, q2);
        depth++;
    }
    const uint32_t n_nodes = level_base;
    uint32_t ms = 0;
    WR_HIP(hipMemcpyAsync(&ms, max_stack, sizeof ms, hipMemcpyDeviceToHost, stream));
    WR_HIP(hipMalloc(&res->nodes, sizeof(Node4) * n_nodes));
    WR_HIP(hipMemcpyAsync(res->nodes, nodes_tmp, sizeof(Node4) * n_nodes, hipMemcpyDeviceToDevice, stream));
    WR_HIP(hipMemcpyAsync(leaf_to_orig, leaf_orig, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, stream));
    WR_HIP(hipStreamSynchronize(stream));
    res->n_nodes = n_nodes;
    res->max_stack = ms;
    res->depth = depth;
    res->iterations = iters;
    res->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return WR_OK;
}

}  // namespace wr
