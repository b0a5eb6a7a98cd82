// wr_traverse.h -- BVH4 traversal with a per-lane LDS stack and the
// watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013), FP32 with an
// FP64 fallback on edge hits.
//
// Semantics follow rtcIntersect1 / rtcOccluded1 as the reference calls them
// (src/accel/embreeaccel.h:336-422): double-sided triangles, hits accepted
// for t in (tmin, tmax], closest hit for extension rays, any hit for shadow
// rays.  Node boxes are tested conservatively (far distance scaled up by
// 1 + 2^-21) so that no watertight hit is lost to rounding in the box test.
#pragma once
#include "wr_device.h"

namespace wr {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;

struct RaySetup {
    float3 o, d;
    float tmin;
    float ix, iy, iz;          // 1/d with zero components replaced by +-1e-20
    int kx, ky, kz;            // watertight axis permutation
    float Sx, Sy, Sz;
    bool nx, ny, nz;           // direction sign per axis (picks near/far slab)
};

__device__ __forceinline__ float pick(float3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

__device__ __forceinline__ RaySetup setup_ray(float3 o, float3 d, float tmin) {
    RaySetup r;
    r.o = o; r.d = d; r.tmin = tmin;
    auto safe = [](float x) { return fabsf(x) < 1e-20f ? copysignf(1e-20f, x) : x; };
    r.ix = 1.0f / safe(d.x); r.iy = 1.0f / safe(d.y); r.iz = 1.0f / safe(d.z);
    r.nx = r.ix < 0.f; r.ny = r.iy < 0.f; r.nz = r.iz < 0.f;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    if (pick(d, kz) < 0.f) { int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    float dz = pick(d, kz);
    r.Sx = pick(d, kx) / dz; r.Sy = pick(d, ky) / dz; r.Sz = 1.0f / dz;
    return r;
}

// FP64 recomputation of the watertight test (edge and vertex hits, where an
// FP32 edge function is exactly zero).
__device__ __forceinline__ double pickd(double x, double y, double z, int k) { return k == 0 ? x : (k == 1 ? y : z); }
__device__ __forceinline__ bool tri_test_fp64(const RaySetup &r, float3 a, float3 b, float3 c, double tmin, double tmax,
                                              double &t_out) {
    const double ox = r.o.x, oy = r.o.y, oz = r.o.z;
    const double ax = a.x - ox, ay = a.y - oy, az = a.z - oz;
    const double bx = b.x - ox, by = b.y - oy, bz = b.z - oz;
    const double cx = c.x - ox, cy = c.y - oy, cz = c.z - oz;
    const double dkx = pickd(r.d.x, r.d.y, r.d.z, r.kx), dky = pickd(r.d.x, r.d.y, r.d.z, r.ky);
    const double dkz = pickd(r.d.x, r.d.y, r.d.z, r.kz);
    const double Sx = dkx / dkz, Sy = dky / dkz, Sz = 1.0 / dkz;
    const double Akz = pickd(ax, ay, az, r.kz), Bkz = pickd(bx, by, bz, r.kz), Ckz = pickd(cx, cy, cz, r.kz);
    const double Ax = pickd(ax, ay, az, r.kx) - Sx * Akz, Ay = pickd(ax, ay, az, r.ky) - Sy * Akz;
    const double Bx = pickd(bx, by, bz, r.kx) - Sx * Bkz, By = pickd(bx, by, bz, r.ky) - Sy * Bkz;
    const double Cx = pickd(cx, cy, cz, r.kx) - Sx * Ckz, Cy = pickd(cx, cy, cz, r.ky) - Sy * Ckz;
    double U, V, W;
    {
        // unfused, as in tri_test: a shared edge gives its two triangles
        // exactly opposite values (the watertightness argument of Woop et al.)
#pragma clang fp contract(off)
        U = Cx * By - Cy * Bx;
        V = Ax * Cy - Ay * Cx;
        W = Bx * Ay - By * Ax;
    }
    if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return false;
    const double det = U + V + W;
    if (det == 0.0) return false;
    const double t = Sz * (U * Akz + V * Bkz + W * Ckz) / det;
    if (!(t > tmin && t <= tmax)) return false;
    t_out = t;
    return true;
}

// Watertight FP32 test; returns true and t for a hit in (tmin, tmax].
__device__ __forceinline__ bool tri_test(const RaySetup &r, float4 va, float4 vb, float4 vc, float tmax, float &t_out) {
    const float3 A = xyz(va) - r.o, B = xyz(vb) - r.o, C = xyz(vc) - r.o;
    const float Akz = pick(A, r.kz), Bkz = pick(B, r.kz), Ckz = pick(C, r.kz);
    const float Ax = fmaf(-r.Sx, Akz, pick(A, r.kx)), Ay = fmaf(-r.Sy, Akz, pick(A, r.ky));
    const float Bx = fmaf(-r.Sx, Bkz, pick(B, r.kx)), By = fmaf(-r.Sy, Bkz, pick(B, r.ky));
    const float Cx = fmaf(-r.Sx, Ckz, pick(C, r.kx)), Cy = fmaf(-r.Sy, Ckz, pick(C, r.ky));
    float U, V, W;
    {
        // The per-vertex projections above may fuse (a vertex then projects to
        // the same value in every triangle that holds it), but the edge
        // functions must not: Cx*By - Cy*Bx with one product fused is not the
        // exact negation of the same edge's value in the neighbouring triangle,
        // Bx*Cy - By*Cx, and a ray through the shared edge could miss both.
#pragma clang fp contract(off)
        U = Cx * By - Cy * Bx;
        V = Ax * Cy - Ay * Cx;
        W = Bx * Ay - By * Ax;
    }
    if (U == 0.f || V == 0.f || W == 0.f) {
        double t;
        if (!tri_test_fp64(r, xyz(va), xyz(vb), xyz(vc), r.tmin, tmax, t)) return false;
        t_out = float(t);
        return true;
    }
    if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
    const float det = U + V + W;
    if (det == 0.f) return false;
    const float T = r.Sz * (U * Akz + V * Bkz + W * Ckz);
    const float t = T / det;
    if (!(t > r.tmin && t <= tmax)) return false;
    t_out = t;
    return true;
}

// Far limit against which a closest-hit traversal culls boxes and stack
// entries: the best t so far widened by 2^-21 (4 ulp).  A triangle that ties
// with the best hit in t may sit in a box whose FP32 entry distance rounds a
// few ulp above that t; culling such a box would make the equal-t rule (lower
// leaf-order index wins) depend on the order in which leaves are reached.
__device__ __forceinline__ float cull_limit(float t) { return fmaf(fabsf(t), 4.76837158e-7f, t); }

// Per-lane traversal stack of (node ref, entry distance).  The first
// kLdsStackDepth entries live in LDS at [k * kBlock + lane] (bank-conflict
// free); deeper entries spill to a per-lane slice of global memory.  The BVH
// builder guarantees at most kStackDepth live entries.
template <int kBlock>
struct LaneStack {
    uint32_t *node;
    float *tnear;
    uint2 *spill;
    int sp;
    __device__ __forceinline__ void push(uint32_t ref, float t) {
        if (sp < kLdsStackDepth) {
            node[sp * kBlock] = ref;
            tnear[sp * kBlock] = t;
        } else {
            spill[sp - kLdsStackDepth] = make_uint2(ref, __float_as_uint(t));
        }
        sp++;
    }
    __device__ __forceinline__ void top(uint32_t &ref, float &t) const {
        if (sp < kLdsStackDepth) {
            ref = node[sp * kBlock];
            t = tnear[sp * kBlock];
        } else {
            const uint2 e = spill[sp - kLdsStackDepth];
            ref = e.x;
            t = __uint_as_float(e.y);
        }
    }
    // node refs only (any-hit traversal never culls a popped entry)
    __device__ __forceinline__ void push_ref(uint32_t ref) {
        if (sp < kLdsStackDepth) node[sp * kBlock] = ref;
        else reinterpret_cast<uint32_t *>(spill + (sp - kLdsStackDepth))[0] = ref;
        sp++;
    }
    __device__ __forceinline__ uint32_t top_ref() const {
        return sp < kLdsStackDepth ? node[sp * kBlock] : reinterpret_cast<const uint32_t *>(spill + (sp - kLdsStackDepth))[0];
    }
};

struct TraversalCounters {
    uint32_t nodes = 0, tris = 0;
};

// Resumable traversal state of one lane.
struct TravState {
    RaySetup r;
    float t;          // current far limit (closest hit so far, or the shadow ray's tmax)
    uint32_t hit;     // leaf-order triangle or kInvalid
    uint32_t cur;     // node reference to process next
};

__device__ __forceinline__ void trav_begin(TravState &ts, float3 o, float3 d, float tmin, float tmax) {
    ts.r = setup_ray(o, d, tmin);
    ts.t = tmax;
    ts.hit = kInvalid;
    ts.cur = 0;  // root (internal node 0)
}

// Processes one node (internal: box tests and pushes; leaf: triangle tests),
// then pops if needed.  Returns true when the lane's ray is finished.
// Closest hit (ANY = false) visits children near-to-far and culls stack entries
// farther than the current hit; any hit (ANY = true) stops at the first hit.
// G is any scene handle with `nodes` and `tris` (DScene, or a per-lane Geo).
template <bool ANY, bool COUNT, int kBlock, class G>
__device__ __forceinline__ bool trav_step(const G &s, TravState &ts, LaneStack<kBlock> &st, TraversalCounters &cnt) {
    const RaySetup &r = ts.r;
    const uint32_t cur = ts.cur;
    if (!(cur & kLeafBit)) {
        const float4 *np = reinterpret_cast<const float4 *>(s.nodes + cur);
        const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
        const uint4 ch = reinterpret_cast<const uint4 *>(np)[6];
        if (COUNT) cnt.nodes++;
        const float ox = r.o.x, oy = r.o.y, oz = r.o.z;
        const float4 nx_ = r.nx ? hx : lx, fx_ = r.nx ? lx : hx;
        const float4 ny_ = r.ny ? hy : ly, fy_ = r.ny ? ly : hy;
        const float4 nz_ = r.nz ? hz : lz, fz_ = r.nz ? lz : hz;
        float tn[4], tf[4];
        tn[0] = fmaxf(fmaxf((nx_.x - ox) * r.ix, (ny_.x - oy) * r.iy), fmaxf((nz_.x - oz) * r.iz, r.tmin));
        tn[1] = fmaxf(fmaxf((nx_.y - ox) * r.ix, (ny_.y - oy) * r.iy), fmaxf((nz_.y - oz) * r.iz, r.tmin));
        tn[2] = fmaxf(fmaxf((nx_.z - ox) * r.ix, (ny_.z - oy) * r.iy), fmaxf((nz_.z - oz) * r.iz, r.tmin));
        tn[3] = fmaxf(fmaxf((nx_.w - ox) * r.ix, (ny_.w - oy) * r.iy), fmaxf((nz_.w - oz) * r.iz, r.tmin));
        tf[0] = fminf(fminf((fx_.x - ox) * r.ix, (fy_.x - oy) * r.iy), (fz_.x - oz) * r.iz);
        tf[1] = fminf(fminf((fx_.y - ox) * r.ix, (fy_.y - oy) * r.iy), (fz_.y - oz) * r.iz);
        tf[2] = fminf(fminf((fx_.z - ox) * r.ix, (fy_.z - oy) * r.iy), (fz_.z - oz) * r.iz);
        tf[3] = fminf(fminf((fx_.w - ox) * r.ix, (fy_.w - oy) * r.iy), (fz_.w - oz) * r.iz);
        uint32_t c[4] = {ch.x, ch.y, ch.z, ch.w};
        float key[4];
        const float tc = ANY ? ts.t : cull_limit(ts.t);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const bool h = c[i] != kEmptyChild && tn[i] <= fminf(tf[i] * 1.00000048f, tc);
            key[i] = h ? tn[i] : INFINITY;
        }
        if (!ANY) {
            // sort (key, child) ascending: 5-comparator network
#define WR_CSWAP(a, b)                                                          \
    if (key[b] < key[a]) {                                                      \
        float tk = key[a]; key[a] = key[b]; key[b] = tk;                        \
        uint32_t tc = c[a]; c[a] = c[b]; c[b] = tc;                             \
    }
            WR_CSWAP(0, 1) WR_CSWAP(2, 3) WR_CSWAP(0, 2) WR_CSWAP(1, 3) WR_CSWAP(1, 2)
#undef WR_CSWAP
        }
        // push hits farthest first; continue with the nearest
        uint32_t next = kInvalid;
        float next_t = 0.f;
#pragma unroll
        for (int i = 3; i >= 0; i--) {
            if (key[i] != INFINITY) {
                if (next != kInvalid) {
                    if (ANY) st.push_ref(next);
                    else st.push(next, next_t);
                }
                next = c[i];
                next_t = key[i];
            }
        }
        if (next != kInvalid) { ts.cur = next; return false; }
    } else {
        const uint32_t first = cur & 0x0FFFFFFFu;
        const uint32_t count = ((cur >> 28) & 0x7u) + 1u;
        if (COUNT) cnt.tris += count;
        for (uint32_t k = 0; k < count; k++) {
            const float4 *tp = s.tris + 3 * size_t(first + k);
            float t;
            // at equal t the lower leaf-order index wins (kInvalid is the largest), so
            // the closest hit does not depend on the order in which leaves are tested
            if (tri_test(r, tp[0], tp[1], tp[2], ts.t, t) && (ANY || t < ts.t || first + k < ts.hit)) {
                ts.t = t;
                ts.hit = first + k;
                if (ANY) return true;
            }
        }
    }
    // pop, skipping entries farther than the current hit (any hit: the far
    // limit has not moved since the push)
    if (ANY) {
        if (st.sp == 0) return true;
        st.sp--;
        ts.cur = st.top_ref();
        return false;
    }
    for (;;) {
        if (st.sp == 0) return true;
        st.sp--;
        uint32_t ref;
        float tn;
        st.top(ref, tn);
        if (tn <= cull_limit(ts.t)) { ts.cur = ref; return false; }
    }
}

// Closest hit, one speculative while-while round (Aila and Laine, "Understanding
// the Efficiency of Ray Traversal on GPUs", HPG 2009): each lane descends inner
// nodes and postpones the first leaf it meets, continuing speculatively with
// the next node, until every lane of the wave holds a leaf or has finished (or
// kSpecSteps steps have run); then the wave tests its postponed leaves
// together, so the box-test and triangle-test code each run with dense lanes
// (one node per step with both branches left a third of the lanes active).
// Speculation visits ~3% more nodes; short rounds keep that small: 4 steps
// were fastest of 2..32 (extension 12-18% faster on the 1.31M-triangle and
// gap-room frames, identical hits).  The push/pop order is the one-node-per-
// step order, so the builders' stack bound holds.  All lanes of the wave must
// call it (it takes a ballot); `live` marks lanes with a ray.  Returns true
// when the lane's ray is finished.
constexpr int kSpecSteps = 4;
template <bool COUNT, int kBlock, class G>
__device__ __forceinline__ bool trav_round(const G &s, TravState &ts, LaneStack<kBlock> &st, TraversalCounters &cnt,
                                           bool live) {
    const RaySetup &r = ts.r;
    uint32_t leaf = kInvalid;  // postponed leaf
    bool more = live;          // ts.cur is a node still to process
    for (int it = 0; it < kSpecSteps; it++) {
        // the wave stops once every lane holds a leaf or has no node left
        if (__ballot(more && leaf == kInvalid) == 0) break;
        const uint32_t cur = ts.cur;
        if (!more || ((cur & kLeafBit) && leaf != kInvalid)) continue;  // waits with its leaf
        bool pop = false;
        if (cur & kLeafBit) {
            leaf = cur;
            pop = true;
        } else {
            const float4 *np = reinterpret_cast<const float4 *>(s.nodes + cur);
            const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
            const uint4 ch = reinterpret_cast<const uint4 *>(np)[6];
            if (COUNT) cnt.nodes++;
            const float ox = r.o.x, oy = r.o.y, oz = r.o.z;
            const float4 nx_ = r.nx ? hx : lx, fx_ = r.nx ? lx : hx;
            const float4 ny_ = r.ny ? hy : ly, fy_ = r.ny ? ly : hy;
            const float4 nz_ = r.nz ? hz : lz, fz_ = r.nz ? lz : hz;
            float tn[4], tf[4];
            tn[0] = fmaxf(fmaxf((nx_.x - ox) * r.ix, (ny_.x - oy) * r.iy), fmaxf((nz_.x - oz) * r.iz, r.tmin));
            tn[1] = fmaxf(fmaxf((nx_.y - ox) * r.ix, (ny_.y - oy) * r.iy), fmaxf((nz_.y - oz) * r.iz, r.tmin));
            tn[2] = fmaxf(fmaxf((nx_.z - ox) * r.ix, (ny_.z - oy) * r.iy), fmaxf((nz_.z - oz) * r.iz, r.tmin));
            tn[3] = fmaxf(fmaxf((nx_.w - ox) * r.ix, (ny_.w - oy) * r.iy), fmaxf((nz_.w - oz) * r.iz, r.tmin));
            tf[0] = fminf(fminf((fx_.x - ox) * r.ix, (fy_.x - oy) * r.iy), (fz_.x - oz) * r.iz);
            tf[1] = fminf(fminf((fx_.y - ox) * r.ix, (fy_.y - oy) * r.iy), (fz_.y - oz) * r.iz);
            tf[2] = fminf(fminf((fx_.z - ox) * r.ix, (fy_.z - oy) * r.iy), (fz_.z - oz) * r.iz);
            tf[3] = fminf(fminf((fx_.w - ox) * r.ix, (fy_.w - oy) * r.iy), (fz_.w - oz) * r.iz);
            uint32_t c[4] = {ch.x, ch.y, ch.z, ch.w};
            float key[4];
            const float tc = cull_limit(ts.t);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const bool h = c[i] != kEmptyChild && tn[i] <= fminf(tf[i] * 1.00000048f, tc);
                key[i] = h ? tn[i] : INFINITY;
            }
#define WR_CSWAP(a, b)                                                          \
    if (key[b] < key[a]) {                                                      \
        float tk = key[a]; key[a] = key[b]; key[b] = tk;                        \
        uint32_t tc = c[a]; c[a] = c[b]; c[b] = tc;                             \
    }
            WR_CSWAP(0, 1) WR_CSWAP(2, 3) WR_CSWAP(0, 2) WR_CSWAP(1, 3) WR_CSWAP(1, 2)
#undef WR_CSWAP
            uint32_t next = kInvalid;
            float next_t = 0.f;
#pragma unroll
            for (int i = 3; i >= 0; i--) {
                if (key[i] != INFINITY) {
                    if (next != kInvalid) st.push(next, next_t);
                    next = c[i];
                    next_t = key[i];
                }
            }
            if (next != kInvalid) ts.cur = next;
            else pop = true;
        }
        if (pop) {
            more = false;
            while (st.sp > 0) {
                st.sp--;
                uint32_t ref;
                float tn;
                st.top(ref, tn);
                if (tn <= cull_limit(ts.t)) {
                    ts.cur = ref;
                    more = true;
                    break;
                }
            }
        }
    }
    if (leaf != kInvalid) {
        const uint32_t first = leaf & 0x0FFFFFFFu;
        const uint32_t count = ((leaf >> 28) & 0x7u) + 1u;
        if (COUNT) cnt.tris += count;
        for (uint32_t k = 0; k < count; k++) {
            const float4 *tp = s.tris + 3 * size_t(first + k);
            float t;
            if (tri_test(r, tp[0], tp[1], tp[2], ts.t, t) && (t < ts.t || first + k < ts.hit)) {  // ties: lower index
                ts.t = t;
                ts.hit = first + k;
            }
        }
    }
    return live && !more;
}

// Any hit, the same round (10% faster shadow rays on the 1.31M-triangle frame):
// children in slot order (no sort), a stack of node refs only, and the
// postponed leaf ends the ray at its first hit.
template <bool COUNT, int kBlock, class G>
__device__ __forceinline__ bool trav_round_any(const G &s, TravState &ts, LaneStack<kBlock> &st, TraversalCounters &cnt,
                                               bool live) {
    const RaySetup &r = ts.r;
    uint32_t leaf = kInvalid;
    bool more = live;
    for (int it = 0; it < kSpecSteps; it++) {
        if (__ballot(more && leaf == kInvalid) == 0) break;
        const uint32_t cur = ts.cur;
        if (!more || ((cur & kLeafBit) && leaf != kInvalid)) continue;
        bool pop = false;
        if (cur & kLeafBit) {
            leaf = cur;
            pop = true;
        } else {
            const float4 *np = reinterpret_cast<const float4 *>(s.nodes + cur);
            const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
            const uint4 ch = reinterpret_cast<const uint4 *>(np)[6];
            if (COUNT) cnt.nodes++;
            const float ox = r.o.x, oy = r.o.y, oz = r.o.z;
            const float4 nx_ = r.nx ? hx : lx, fx_ = r.nx ? lx : hx;
            const float4 ny_ = r.ny ? hy : ly, fy_ = r.ny ? ly : hy;
            const float4 nz_ = r.nz ? hz : lz, fz_ = r.nz ? lz : hz;
            float tn[4], tf[4];
            tn[0] = fmaxf(fmaxf((nx_.x - ox) * r.ix, (ny_.x - oy) * r.iy), fmaxf((nz_.x - oz) * r.iz, r.tmin));
            tn[1] = fmaxf(fmaxf((nx_.y - ox) * r.ix, (ny_.y - oy) * r.iy), fmaxf((nz_.y - oz) * r.iz, r.tmin));
            tn[2] = fmaxf(fmaxf((nx_.z - ox) * r.ix, (ny_.z - oy) * r.iy), fmaxf((nz_.z - oz) * r.iz, r.tmin));
            tn[3] = fmaxf(fmaxf((nx_.w - ox) * r.ix, (ny_.w - oy) * r.iy), fmaxf((nz_.w - oz) * r.iz, r.tmin));
            tf[0] = fminf(fminf((fx_.x - ox) * r.ix, (fy_.x - oy) * r.iy), (fz_.x - oz) * r.iz);
            tf[1] = fminf(fminf((fx_.y - ox) * r.ix, (fy_.y - oy) * r.iy), (fz_.y - oz) * r.iz);
            tf[2] = fminf(fminf((fx_.z - ox) * r.ix, (fy_.z - oy) * r.iy), (fz_.z - oz) * r.iz);
            tf[3] = fminf(fminf((fx_.w - ox) * r.ix, (fy_.w - oy) * r.iy), (fz_.w - oz) * r.iz);
            const uint32_t c[4] = {ch.x, ch.y, ch.z, ch.w};
            uint32_t next = kInvalid;
#pragma unroll
            for (int i = 3; i >= 0; i--) {
                if (c[i] != kEmptyChild && tn[i] <= fminf(tf[i] * 1.00000048f, ts.t)) {
                    if (next != kInvalid) st.push_ref(next);
                    next = c[i];
                }
            }
            if (next != kInvalid) ts.cur = next;
            else pop = true;
        }
        if (pop) {
            more = st.sp > 0;
            if (more) {
                st.sp--;
                ts.cur = st.top_ref();
            }
        }
    }
    if (leaf != kInvalid) {
        const uint32_t first = leaf & 0x0FFFFFFFu;
        const uint32_t count = ((leaf >> 28) & 0x7u) + 1u;
        if (COUNT) cnt.tris += count;
        for (uint32_t k = 0; k < count; k++) {
            const float4 *tp = s.tris + 3 * size_t(first + k);
            float t;
            if (tri_test(r, tp[0], tp[1], tp[2], ts.t, t)) {
                ts.t = t;
                ts.hit = first + k;
                return live;  // occluded: finished
            }
        }
    }
    return live && !more;
}

// Whole traversal of one ray (used where rays are not refilled per lane).
template <bool ANY, bool COUNT, int kBlock, class G>
__device__ __forceinline__ uint32_t traverse(const G &s, const RaySetup &r, float &t_best, LaneStack<kBlock> &st,
                                             TraversalCounters &cnt) {
    TravState ts;
    ts.r = r;
    ts.t = t_best;
    ts.hit = kInvalid;
    ts.cur = 0;
    st.sp = 0;
    while (!trav_step<ANY, COUNT, kBlock>(s, ts, st, cnt)) {
    }
    t_best = ts.t;
    return ts.hit;
}

// Whole traversal of one ray in speculative rounds (trav_round / trav_round_any):
// the same hits as traverse() (ties in t go to the lower leaf-order index in
// both).  The rounds' ballots count the lanes that are active, so lanes of a
// wave may enter and leave it at different times.
template <bool ANY, int kBlock, class G>
__device__ __forceinline__ uint32_t traverse_rounds(const G &s, const RaySetup &r, float &t_best, LaneStack<kBlock> &st,
                                                    TraversalCounters &cnt) {
    TravState ts;
    ts.r = r;
    ts.t = t_best;
    ts.hit = kInvalid;
    ts.cur = 0;
    st.sp = 0;
    bool done;
    do {
        done = ANY ? trav_round_any<false, kBlock>(s, ts, st, cnt, true) : trav_round<false, kBlock>(s, ts, st, cnt, true);
    } while (!done);
    t_best = ts.t;
    return ts.hit;
}

}  // namespace wr
