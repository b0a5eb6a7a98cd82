// wr_bvh.cpp -- host BVH build: binned SAH over triangle bounds (BVH2), then
// collapsed to the 4-wide, 128-byte-node layout of wr_layout.h.
//
// Replaces rtcCommitScene's BVH8/Triangle4 SAH build (reference:
// src/accel/embreeaccel.h:466-469; Embree kernels/common/scene.cpp:120-140).
// The hit semantics do not depend on the tree: any BVH gives the same closest
// hit, so the tree shape is a performance choice only.
#include "wr_layout.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <future>
#include <new>
#include <vector>

namespace wr {
namespace {

constexpr int kBins = 32;
constexpr float kTraversalCost = 1.0f;
constexpr float kTriangleCost = 1.0f;
constexpr uint32_t kParallelThreshold = 65536;

struct Ref {
    float lo[3], hi[3], c[3];
    uint32_t tri;
};

struct B2 {
    float lo[3], hi[3];
    uint32_t left;         // internal: index of left child (right = left + 1)
    uint32_t first, count; // leaf range in refs (count > 0)
};

struct Aabb {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const float *l, const float *h) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], l[k]);
            hi[k] = std::max(hi[k], h[k]);
        }
    }
    void grow(const Aabb &b) { grow(b.lo, b.hi); }
    float half_area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.f;
        return dx * dy + dy * dz + dz * dx;
    }
};

struct Builder {
    std::vector<Ref> refs;
    std::vector<B2> nodes;
    std::atomic<uint32_t> n_nodes{0};

    uint32_t alloc(uint32_t k) { return n_nodes.fetch_add(k); }

    void make_leaf(uint32_t ni, uint32_t begin, uint32_t end, const Aabb &box) {
        B2 &n = nodes[ni];
        std::memcpy(n.lo, box.lo, sizeof n.lo);
        std::memcpy(n.hi, box.hi, sizeof n.hi);
        n.left = 0;
        n.first = begin;
        n.count = end - begin;
    }

    void build(uint32_t ni, uint32_t begin, uint32_t end, int depth) {
        Aabb box, cbox;
        for (uint32_t i = begin; i < end; i++) {
            box.grow(refs[i].lo, refs[i].hi);
            cbox.grow(refs[i].c, refs[i].c);
        }
        const uint32_t n = end - begin;
        if (n == 1) { make_leaf(ni, begin, end, box); return; }
        // binned SAH
        float best_cost = FLT_MAX;
        int best_axis = -1, best_split = -1;
        for (int k = 0; k < 3; k++) {
            float ext = cbox.hi[k] - cbox.lo[k];
            if (!(ext > 0.f)) continue;
            float scale = kBins / ext;
            Aabb bb[kBins];
            uint32_t bc[kBins] = {0};
            for (uint32_t i = begin; i < end; i++) {
                int b = std::min(kBins - 1, std::max(0, int((refs[i].c[k] - cbox.lo[k]) * scale)));
                bc[b]++;
                bb[b].grow(refs[i].lo, refs[i].hi);
            }
            float right_area[kBins];
            uint32_t right_count[kBins];
            Aabb acc;
            uint32_t cnt = 0;
            for (int b = kBins - 1; b > 0; b--) {
                acc.grow(bb[b]);
                cnt += bc[b];
                right_area[b] = acc.half_area();
                right_count[b] = cnt;
            }
            Aabb lacc;
            uint32_t lcnt = 0;
            for (int b = 0; b < kBins - 1; b++) {
                lacc.grow(bb[b]);
                lcnt += bc[b];
                if (lcnt == 0 || right_count[b + 1] == 0) continue;
                float cost = lacc.half_area() * lcnt + right_area[b + 1] * right_count[b + 1];
                if (cost < best_cost) { best_cost = cost; best_axis = k; best_split = b; }
            }
        }
        const float parent_area = box.half_area();
        const float leaf_cost = kTriangleCost * n;
        const float split_cost = parent_area > 0.f ? kTraversalCost + kTriangleCost * best_cost / parent_area : FLT_MAX;
        if (n <= (uint32_t)kMaxLeafTris && (best_axis < 0 || leaf_cost <= split_cost)) {
            make_leaf(ni, begin, end, box);
            return;
        }
        uint32_t mid;
        if (best_axis < 0) {
            mid = begin + n / 2;  // coincident centroids: split by count
        } else {
            const int k = best_axis;
            const float scale = kBins / (cbox.hi[k] - cbox.lo[k]);
            auto it = std::partition(refs.begin() + begin, refs.begin() + end, [&](const Ref &r) {
                int b = std::min(kBins - 1, std::max(0, int((r.c[k] - cbox.lo[k]) * scale)));
                return b <= best_split;
            });
            mid = uint32_t(it - refs.begin());
            if (mid == begin || mid == end) mid = begin + n / 2;
        }
        const uint32_t l = alloc(2);
        B2 &nd = nodes[ni];
        std::memcpy(nd.lo, box.lo, sizeof nd.lo);
        std::memcpy(nd.hi, box.hi, sizeof nd.hi);
        nd.left = l;
        nd.first = 0;
        nd.count = 0;
        if (n > kParallelThreshold && depth < 8) {
            auto fut = std::async(std::launch::async, [=] { build(l, begin, mid, depth + 1); });
            build(l + 1, mid, end, depth + 1);
            fut.get();
        } else {
            build(l, begin, mid, depth + 1);
            build(l + 1, mid, end, depth + 1);
        }
    }
};

// Collapse of the binary tree to 4-wide nodes, choosing which binary nodes
// become the children of each wide node by the SAH (Ylitie, Karras and Laine,
// "Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs",
// HPG 2017, section 4.1): cost[n][i-1] is the least SAH cost of covering the
// subtree of binary node n with at most i wide-node children, computed bottom
// up; binary leaves stay whole, and an inner binary node of at most
// kMaxLeafTris triangles may become one wide leaf.
struct Collapser {
    const std::vector<B2> &b2;
    std::vector<Node4> out;
    uint32_t max_stack = 0, max_depth = 0;
    std::vector<float> cost;       // 4 per binary node
    std::vector<uint8_t> split;    // 4 per binary node: [0] 1 = one wide leaf; [i-1] (i > 1): 0 = as with i-1
                                   // children, k = k children from the left subtree and i-k from the right
    std::vector<uint8_t> top_split;  // per inner binary node: slots of its own wide node given to the left
    std::vector<uint32_t> cnt, first;  // triangles under a binary node, first of them in leaf order

    explicit Collapser(const std::vector<B2> &b) : b2(b) {}

    static float area(const B2 &n) {
        float dx = n.hi[0] - n.lo[0], dy = n.hi[1] - n.lo[1], dz = n.hi[2] - n.lo[2];
        return dx * dy + dy * dz + dz * dx;
    }

    // children always have larger ids than their parent (allocated after it)
    void optimise() {
        const size_t n = b2.size();
        cost.assign(4 * n, 0.f);
        split.assign(4 * n, 0);
        cnt.assign(n, 0);
        first.assign(n, 0);
        top_split.assign(n, 0);
        for (size_t q = n; q-- > 0;) {
            const B2 &b = b2[q];
            const float A = area(b);
            float *c = &cost[4 * q];
            uint8_t *sp = &split[4 * q];
            if (b.count > 0) {
                cnt[q] = b.count;
                first[q] = b.first;
                for (int i = 0; i < 4; i++) c[i] = kTriangleCost * A * float(b.count);
                sp[0] = 1;
                continue;
            }
            const uint32_t l = b.left, r = l + 1;
            cnt[q] = cnt[l] + cnt[r];
            first[q] = first[l];  // the left subtree's triangles come first
            float dist[5];
            uint8_t dk[5];
            for (int j = 2; j <= 4; j++) {  // k = 1 first: a valid split even if a cost is not finite
                dist[j] = cost[4 * l] + cost[4 * r + j - 2];
                dk[j] = 1;
                for (int k = 2; k < j; k++) {
                    const float v = cost[4 * l + k - 1] + cost[4 * r + j - k - 1];
                    if (v < dist[j]) { dist[j] = v; dk[j] = uint8_t(k); }
                }
            }
            const float inner = kTraversalCost * A + dist[4];
            const float leaf = cnt[q] <= uint32_t(kMaxLeafTris) ? kTriangleCost * A * float(cnt[q]) : FLT_MAX;
            c[0] = std::min(inner, leaf);
            sp[0] = leaf <= inner ? 1 : 0;
            for (int i = 2; i <= 4; i++) {
                if (dist[i] < c[i - 2]) { c[i - 1] = dist[i]; sp[i - 1] = dk[i]; }
                else { c[i - 1] = c[i - 2]; sp[i - 1] = 0; }
            }
            top_split[q] = dk[4];
        }
    }
    // Appends the wide children that cover binary node m with at most i slots.
    void expand(uint32_t m, int i, uint32_t *kids, int &k) const {
        for (;;) {
            const int s = i > 1 ? split[4 * size_t(m) + i - 1] : 0;
            if (i == 1 || b2[m].count > 0) { kids[k++] = m; return; }
            if (s == 0) { i--; continue; }
            expand(b2[m].left, s, kids, k);
            expand(b2[m].left + 1, i - s, kids, k);
            return;
        }
    }
    bool wide_leaf(uint32_t m) const { return split[4 * size_t(m)] != 0; }

    // stack_before: entries that may sit on the stack when this node is visited
    uint32_t collapse(uint32_t q, uint32_t stack_before, uint32_t depth) {
        uint32_t kids[4];
        int k = 0;
        if (b2[q].count > 0) {
            kids[k++] = q;  // a root that is one leaf
        } else {
            const int s = top_split[q];
            expand(b2[q].left, s, kids, k);
            expand(b2[q].left + 1, 4 - s, kids, k);
        }
        // slot order = decreasing area: any-hit traversal visits the slots in
        // order, and a larger box is likelier to hold an occluder
        std::stable_sort(kids, kids + k, [&](uint32_t a, uint32_t b) { return area(b2[a]) > area(b2[b]); });
        const uint32_t me = uint32_t(out.size());
        out.emplace_back();
        Node4 nd;
        for (int i = 0; i < 4; i++) {
            nd.lo_x[i] = nd.lo_y[i] = nd.lo_z[i] = FLT_MAX;
            nd.hi_x[i] = nd.hi_y[i] = nd.hi_z[i] = -FLT_MAX;
            nd.child[i] = kEmptyChild;
            nd.count[i] = 0;
        }
        // worst case: all k children hit; k-1 stay on the stack while descending
        const uint32_t stack_here = stack_before + uint32_t(k);
        max_stack = std::max(max_stack, stack_here);
        max_depth = std::max(max_depth, depth + 1);
        for (int i = 0; i < k; i++) {
            const B2 &c = b2[kids[i]];
            nd.lo_x[i] = c.lo[0]; nd.lo_y[i] = c.lo[1]; nd.lo_z[i] = c.lo[2];
            nd.hi_x[i] = c.hi[0]; nd.hi_y[i] = c.hi[1]; nd.hi_z[i] = c.hi[2];
            if (wide_leaf(kids[i])) {
                nd.child[i] = kLeafBit | ((cnt[kids[i]] - 1u) << 28) | first[kids[i]];
                nd.count[i] = cnt[kids[i]];
            } else {
                nd.child[i] = collapse(kids[i], stack_before + uint32_t(k) - 1, depth + 1);
                nd.count[i] = 0;
            }
        }
        out[me] = nd;
        return me;
    }
};

}  // namespace

bool build_bvh4(const float *verts, uint32_t nv, const uint32_t *tris, uint32_t nt, const uint32_t *tri_geom,
                BvhBuildResult *res) {
    *res = BvhBuildResult{};
    Builder b;
    try {
        b.refs.resize(nt);
        b.nodes.resize(nt > 0 ? 2 * size_t(nt) - 1 : 1);
    } catch (const std::bad_alloc &) {
        return false;
    }
    for (uint32_t t = 0; t < nt; t++) {
        Ref &r = b.refs[t];
        for (int k = 0; k < 3; k++) { r.lo[k] = FLT_MAX; r.hi[k] = -FLT_MAX; }
        for (int j = 0; j < 3; j++) {
            const float *p = verts + 3 * size_t(tris[3 * size_t(t) + j]);
            for (int k = 0; k < 3; k++) {
                r.lo[k] = std::min(r.lo[k], p[k]);
                r.hi[k] = std::max(r.hi[k], p[k]);
            }
        }
        for (int k = 0; k < 3; k++) r.c[k] = 0.5f * (r.lo[k] + r.hi[k]);
        r.tri = t;
    }
    std::vector<Node4> nodes;
    uint32_t max_stack = 1, depth = 1;
    if (nt == 0) {
        Node4 nd;
        for (int i = 0; i < 4; i++) {
            nd.lo_x[i] = nd.lo_y[i] = nd.lo_z[i] = FLT_MAX;
            nd.hi_x[i] = nd.hi_y[i] = nd.hi_z[i] = -FLT_MAX;
            nd.child[i] = kEmptyChild;
            nd.count[i] = 0;
        }
        nodes.push_back(nd);
    } else {
        const uint32_t root = b.alloc(1);
        b.build(root, 0, nt, 0);
        b.nodes.resize(b.n_nodes.load());
        Collapser c(b.nodes);
        c.out.reserve(b.nodes.size() / 3 + 1);
        c.optimise();
        c.collapse(root, 0, 0);
        nodes.swap(c.out);
        max_stack = c.max_stack;
        depth = c.max_depth;
    }
    res->n_nodes = uint32_t(nodes.size());
    res->nodes = static_cast<Node4 *>(std::aligned_alloc(128, sizeof(Node4) * nodes.size()));
    const size_t tri_bytes = (sizeof(TriRecord) * std::max<size_t>(nt, 1) + 127) / 128 * 128;
    res->tris = static_cast<TriRecord *>(std::aligned_alloc(128, tri_bytes));
    if (!res->nodes || !res->tris) { free_bvh(res); return false; }
    std::memcpy(res->nodes, nodes.data(), sizeof(Node4) * nodes.size());
    for (uint32_t i = 0; i < nt; i++) {
        const uint32_t t = b.refs[i].tri;
        TriRecord &r = res->tris[i];
        const float *p0 = verts + 3 * size_t(tris[3 * size_t(t) + 0]);
        const float *p1 = verts + 3 * size_t(tris[3 * size_t(t) + 1]);
        const float *p2 = verts + 3 * size_t(tris[3 * size_t(t) + 2]);
        for (int k = 0; k < 3; k++) { r.v0[k] = p0[k]; r.v1[k] = p1[k]; r.v2[k] = p2[k]; }
        r.orig = t;
        r.geom = tri_geom[t];
        r.pad = 0;
    }
    res->n_tris = nt;
    res->max_stack = max_stack;
    res->depth = depth;
    return true;
}

const char *verify_bvh4(const BvhBuildResult &r, const float *verts, const uint32_t *tris, uint32_t nt) {
    if (r.n_tris != nt) return "leaf triangle count differs from the input";
    std::vector<uint8_t> seen(nt, 0);
    for (uint32_t i = 0; i < nt; i++) {
        const uint32_t o = r.tris[i].orig;
        if (o >= nt || seen[o]) return "triangle missing or duplicated in leaf order";
        seen[o] = 1;
        const float *p[3] = {verts + 3 * size_t(tris[3 * size_t(o)]), verts + 3 * size_t(tris[3 * size_t(o) + 1]),
                             verts + 3 * size_t(tris[3 * size_t(o) + 2])};
        const float *q[3] = {r.tris[i].v0, r.tris[i].v1, r.tris[i].v2};
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++)
                if (p[j][k] != q[j][k]) return "leaf triangle vertices differ from the input";
    }
    std::vector<uint32_t> leaf_refs(nt, 0);
    std::vector<uint32_t> node_refs(r.n_nodes, 0);
    for (uint32_t n = 0; n < r.n_nodes; n++) {
        const Node4 &nd = r.nodes[n];
        for (int c = 0; c < 4; c++) {
            const uint32_t ch = nd.child[c];
            if (ch == kEmptyChild) continue;
            const float lo[3] = {nd.lo_x[c], nd.lo_y[c], nd.lo_z[c]}, hi[3] = {nd.hi_x[c], nd.hi_y[c], nd.hi_z[c]};
            if (ch & kLeafBit) {
                const uint32_t first = ch & 0x0FFFFFFFu, cnt = ((ch >> 28) & 7u) + 1u;
                if (cnt != nd.count[c] || first + cnt > nt) return "bad leaf reference";
                for (uint32_t i = first; i < first + cnt; i++) {
                    leaf_refs[i]++;
                    const float *q[3] = {r.tris[i].v0, r.tris[i].v1, r.tris[i].v2};
                    for (int j = 0; j < 3; j++)
                        for (int k = 0; k < 3; k++)
                            if (q[j][k] < lo[k] || q[j][k] > hi[k]) return "leaf box does not enclose its triangle";
                }
            } else {
                if (ch >= r.n_nodes || ch <= n) return "bad child node index";
                node_refs[ch]++;
                const Node4 &cn = r.nodes[ch];
                for (int cc = 0; cc < 4; cc++) {
                    if (cn.child[cc] == kEmptyChild) continue;
                    const float clo[3] = {cn.lo_x[cc], cn.lo_y[cc], cn.lo_z[cc]};
                    const float chi[3] = {cn.hi_x[cc], cn.hi_y[cc], cn.hi_z[cc]};
                    for (int k = 0; k < 3; k++)
                        if (clo[k] < lo[k] || chi[k] > hi[k]) return "child box not enclosed by its parent entry";
                }
            }
        }
    }
    for (uint32_t i = 0; i < nt; i++)
        if (leaf_refs[i] != 1) return "triangle referenced by zero or several leaves";
    for (uint32_t n = 1; n < r.n_nodes; n++)
        if (node_refs[n] != 1) return "node referenced by zero or several parents";
    if (r.max_stack > uint32_t(kStackDepth)) return "stack bound exceeds the traversal stack";
    return "";
}

void free_bvh(BvhBuildResult *r) {
    std::free(r->nodes);
    std::free(r->tris);
    r->nodes = nullptr;
    r->tris = nullptr;
}

}  // namespace wr
