// wr_traverse.h -- BVH4 traversal with a per-lane LDS stack and the
// watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013), FP32 with the
// paper's FP64 recomputation of the edge functions on edge hits.
//
// Semantics follow rtcIntersect1 / rtcOccluded1 as the reference calls them
// (src/accel/embreeaccel.h:336-422): double-sided triangles, hits accepted
// for t in (tmin, tmax], closest hit for extension rays, any hit for shadow
// rays.  Node boxes are tested conservatively (exit distances widened by the
// bound on the slab arithmetic's rounding error) so that no watertight hit is
// lost to rounding in the box test.
#pragma once
#include "wr_device.h"

namespace wr {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;

struct RaySetup {
    float3 o;
    float tmin;
    float ix, iy, iz;          // 1/d with zero components replaced by +-1e-20
    float nnx, nny, nnz;       // per axis -(o/d), lowered (near planes) and raised (far planes) by
    float nfx, nfy, nfz;       // 2^-22 |o/d|: a plane p is crossed at fma(p, 1/d, nn) or fma(p, 1/d, nf)
    int kx, ky, kz;            // watertight axis permutation
    float Sx, Sy, Sz;
    uint32_t onx, ony, onz;    // byte offsets of the near-plane rows in a Node4 (the far row is at offset ^ 16)
};

__device__ __forceinline__ float pick(float3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// Reciprocals use the hardware v_rcp_f32 (about 1 ulp): box exit distances are
// widened by 4 ulp, and the watertight shear only has to be the same for every
// triangle a ray meets, not correctly rounded.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ RaySetup setup_ray(float3 o, float3 d, float tmin) {
    RaySetup r;
    r.o = o; r.tmin = tmin;
    auto safe = [](float x) { return fabsf(x) < 1e-20f ? copysignf(1e-20f, x) : x; };
    r.ix = rcp(safe(d.x)); r.iy = rcp(safe(d.y)); r.iz = rcp(safe(d.z));
    const float mx = o.x * r.ix, my = o.y * r.iy, mz = o.z * r.iz;
    const float ex = 2.38418579e-7f * fabsf(mx), ey = 2.38418579e-7f * fabsf(my), ez = 2.38418579e-7f * fabsf(mz);
    r.nnx = -mx - ex; r.nny = -my - ey; r.nnz = -mz - ez;
    r.nfx = -mx + ex; r.nfy = -my + ey; r.nfz = -mz + ez;
    // a Node4 holds rows lo_x, hi_x, lo_y, hi_y, lo_z, hi_z of 16 bytes: the near
    // plane of an axis is its lo row for a positive direction, its hi row otherwise
    r.onx = r.ix < 0.f ? 16u : 0u;
    r.ony = r.iy < 0.f ? 48u : 32u;
    r.onz = r.iz < 0.f ? 80u : 64u;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    if (pick(d, kz) < 0.f) { int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sz = rcp(pick(d, kz));
    r.Sx = pick(d, kx) * r.Sz; r.Sy = pick(d, ky) * r.Sz;
    return r;
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
        // An edge function rounded to zero: recompute the three in FP64 from the
        // same projected coordinates (Woop et al., section 3.3).  A product of
        // two floats is exact in FP64, so each difference has the exact sign,
        // and a shared edge still gives its two triangles opposite values.
        double Ud, Vd, Wd;
        {
#pragma clang fp contract(off)
            Ud = double(Cx) * double(By) - double(Cy) * double(Bx);
            Vd = double(Ax) * double(Cy) - double(Ay) * double(Cx);
            Wd = double(Bx) * double(Ay) - double(By) * double(Ax);
        }
        if ((Ud < 0.0 || Vd < 0.0 || Wd < 0.0) && (Ud > 0.0 || Vd > 0.0 || Wd > 0.0)) return false;
        U = float(Ud);
        V = float(Vd);
        W = float(Wd);
    } else if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) {
        return false;
    }
    const float det = U + V + W;
    if (det == 0.f) return false;
    const float T = r.Sz * (U * Akz + V * Bkz + W * Ckz);
    const float t = T * rcp(det);
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

// Per-lane traversal stack.  Closest hit keeps 64-bit entries (node ref in
// the low half, entry distance bits in the high half), any hit node refs only.
// The first kLdsStackDepth entries live in LDS at [k * kBlock + lane]
// (bank-conflict free: one 64-bit or 32-bit LDS access per push and pop);
// deeper entries spill to the block's slice of a global region, at
// [(k - kLdsStackDepth) * kBlock + lane] (a wave's spills coalesce, and the
// address is a scalar base plus a 32-bit lane offset).  The pointers carry
// their address spaces, so the two branches of a pop stay an LDS load and a
// global load (as generic pointers the compiler merged them into one flat load
// and split the entry into two dependent loads).  The BVH builder guarantees
// at most kStackDepth live entries.
typedef __attribute__((address_space(3))) uint64_t LdsU64;
typedef __attribute__((address_space(3))) uint32_t LdsU32;
typedef __attribute__((address_space(1))) uint64_t GlobalU64;
typedef __attribute__((address_space(1))) uint32_t GlobalU32;
template <int kBlock>
struct LaneStack {
    LdsU64 *ent;        // closest hit: LDS entries of this lane
    LdsU32 *node;       // any hit: LDS node refs of this lane
    GlobalU64 *spill;   // this block's spill slice (kSpillDepth * kBlock entries)
    int sp;
    // spill_region: kSpillDepth * kBlock entries per block of the grid
    __device__ __forceinline__ LaneStack(uint64_t *lds_ent, uint32_t *lds_node, uint2 *spill_region)
        : ent((LdsU64 *)lds_ent), node((LdsU32 *)lds_node),
          spill((GlobalU64 *)spill_region + size_t(blockIdx.x) * (kSpillDepth * kBlock)), sp(0) {}
    __device__ __forceinline__ uint32_t spill_index() const { return uint32_t(sp - kLdsStackDepth) * kBlock + __lane_id(); }
    __device__ __forceinline__ void push(uint64_t e) {
        if (sp < kLdsStackDepth) ent[sp * kBlock] = e;
        else spill[spill_index()] = e;
        sp++;
    }
    __device__ __forceinline__ uint64_t top() const {
        return sp < kLdsStackDepth ? ent[sp * kBlock] : spill[spill_index()];
    }
    __device__ __forceinline__ void push_ref(uint32_t ref) {
        if (sp < kLdsStackDepth) node[sp * kBlock] = ref;
        else reinterpret_cast<GlobalU32 *>(spill + spill_index())[0] = ref;
        sp++;
    }
    __device__ __forceinline__ uint32_t top_ref() const {
        return sp < kLdsStackDepth ? node[sp * kBlock] : reinterpret_cast<GlobalU32 *>(spill + spill_index())[0];
    }
};
__device__ __forceinline__ uint64_t stack_entry(uint32_t ref, uint32_t tn_bits) {
    return uint64_t(ref) | (uint64_t(tn_bits) << 32);
}

// Entry and exit distances of the ray through the four child boxes of inner
// node `cur`, and the child refs.  The node is a NodeQ: four 16-byte loads,
// then each plane is decoded to its FP32 value fmaf(q, scale, origin), which
// lies outside the exact child box (the quantiser checks it with the same
// fmaf), so the slab test below sees a box that contains the exact one.  An
// empty slot is rejected by its ref (kEmptyChild).
//
// A plane p is crossed at fma(p, 1/d, -(o/d)): one operation instead of a
// subtraction and a product.  Against the exact (p - o)/d this form errs by
// the rounding of o/d (at most 2^-24 |o/d|, large only along an axis the ray
// hardly moves in) plus a relative part (1/d is v_rcp_f32 to 1 ulp, and the
// final rounding) below 2^-22.  The absolute part is removed per axis: near
// planes use -(o/d) lowered by 2^-22 |o/d| and far planes -(o/d) raised by as
// much, so a computed entry never exceeds the exact one by more than 2^-22 of
// its size, nor a computed exit falls short by more.  A box the ray touches
// (entry <= exit, both >= tmin > 0) then passes entry <= exit * (1 + 2^-20),
// and one entered before the far limit t passes entry <= cull_limit(t).
// Returns the refs; tn holds the entry distances and hit[i] marks the
// children whose box is entered before `far_limit`.
__device__ __forceinline__ uint4 node_test(const NodeQ *nodes, const RaySetup &r, uint32_t cur, float far_limit,
                                           float tn[4], bool hit[4]) {
    // the node's four 16-byte rows
    const uint4 *nb = reinterpret_cast<const uint4 *>(nodes) + 4 * size_t(cur);
    const uint4 r0 = nb[0], r1 = nb[1], r2 = nb[2], r3 = nb[3];
    const float ox = __uint_as_float(r0.x), oy = __uint_as_float(r0.y), oz = __uint_as_float(r0.z);
    const float sx = __uint_as_float(r0.w), sy = __uint_as_float(r3.z), sz = __uint_as_float(r3.w);
    // the near plane of an axis is its low plane for a positive direction, its high plane otherwise
    const uint32_t qnx = r.ix < 0.f ? r1.y : r1.x, qfx = r.ix < 0.f ? r1.x : r1.y;
    const uint32_t qny = r.iy < 0.f ? r1.w : r1.z, qfy = r.iy < 0.f ? r1.z : r1.w;
    const uint32_t qnz = r.iz < 0.f ? r2.y : r2.x, qfz = r.iz < 0.f ? r2.x : r2.y;
    const uint4 ch = make_uint4(r2.z, r2.w, r3.x, r3.y);
    const uint32_t c[4] = {ch.x, ch.y, ch.z, ch.w};
    auto plane = [](uint32_t q, int i, float sc, float org) { return fmaf(float((q >> (8 * i)) & 0xFFu), sc, org); };
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float nx = plane(qnx, i, sx, ox), ny = plane(qny, i, sy, oy), nz = plane(qnz, i, sz, oz);
        const float fx = plane(qfx, i, sx, ox), fy = plane(qfy, i, sy, oy), fz = plane(qfz, i, sz, oz);
        tn[i] = fmaxf(fmaxf(fmaf(nx, r.ix, r.nnx), fmaf(ny, r.iy, r.nny)), fmaxf(fmaf(nz, r.iz, r.nnz), r.tmin));
        const float tf = fminf(fminf(fmaf(fx, r.ix, r.nfx), fmaf(fy, r.iy, r.nfy)), fmaf(fz, r.iz, r.nfz));
        hit[i] = c[i] != kEmptyChild && tn[i] <= fminf(tf * 1.00000095f, far_limit);
    }
    return ch;
}

// The same test on a full-precision Node4: seven loads (the near and far rows
// at the ray's per-axis byte offsets, and the refs), no decoding.  The kernels
// that trace one coherent or register-heavy path per lane (k_visualize,
// k_bdpt) use it, since their traversal is not bound by load requests.  Empty
// slots hold inverted boxes (lo = FLT_MAX, hi = -FLT_MAX), which fail the test.
__device__ __forceinline__ uint4 node_test(const Node4 *nodes, const RaySetup &r, uint32_t cur, float far_limit,
                                           float tn[4], bool hit[4]) {
    const char *nb = reinterpret_cast<const char *>(nodes);
    const uint32_t base = cur * uint32_t(sizeof(Node4));
    auto row = [&](uint32_t off) { return *reinterpret_cast<const float4 *>(nb + (base + off)); };
    const float4 nx_ = row(r.onx), fx_ = row(r.onx ^ 16u);
    const float4 ny_ = row(r.ony), fy_ = row(r.ony ^ 16u);
    const float4 nz_ = row(r.onz), fz_ = row(r.onz ^ 16u);
    const uint4 ch = *reinterpret_cast<const uint4 *>(nb + (base + 96u));
    const float nx[4] = {nx_.x, nx_.y, nx_.z, nx_.w}, fx[4] = {fx_.x, fx_.y, fx_.z, fx_.w};
    const float ny[4] = {ny_.x, ny_.y, ny_.z, ny_.w}, fy[4] = {fy_.x, fy_.y, fy_.z, fy_.w};
    const float nz[4] = {nz_.x, nz_.y, nz_.z, nz_.w}, fz[4] = {fz_.x, fz_.y, fz_.z, fz_.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
        tn[i] = fmaxf(fmaxf(fmaf(nx[i], r.ix, r.nnx), fmaf(ny[i], r.iy, r.nny)), fmaxf(fmaf(nz[i], r.iz, r.nnz), r.tmin));
        const float tf = fminf(fminf(fmaf(fx[i], r.ix, r.nfx), fmaf(fy[i], r.iy, r.nfy)), fmaf(fz[i], r.iz, r.nfz));
        hit[i] = tn[i] <= fminf(tf * 1.00000095f, far_limit);
    }
    return ch;
}
template <class G>
__device__ __forceinline__ uint4 node_test(const G &s, const RaySetup &r, uint32_t cur, float far_limit, float tn[4],
                                           bool hit[4]) {
    return node_test(s.nodes, r, cur, far_limit, tn, hit);
}

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

// Orders the children near to far, pushes the entered ones but the nearest,
// farthest first, and returns the nearest (kInvalid if none is entered).  A
// key is a 64-bit (entry distance bits, child ref) pair, the distance bits
// 0xFFFFFFFF for a child whose box is not entered; entry distances of entered
// boxes are positive, so their bits order as the distances do.  Each of the
// five comparators of the network is one 32-bit compare and four selects, and
// a push is one 64-bit LDS store.
constexpr uint32_t kNoEntry = 0xFFFFFFFFu;
template <int kBlock>
__device__ __forceinline__ uint32_t sort_push(uint2 key[4], LaneStack<kBlock> &st) {
    auto cswap = [&](int a, int b) {
        const bool sw = key[b].y < key[a].y;
        const uint2 lo = sw ? key[b] : key[a], hi = sw ? key[a] : key[b];
        key[a] = lo;
        key[b] = hi;
    };
    cswap(0, 1); cswap(2, 3); cswap(0, 2); cswap(1, 3); cswap(1, 2);
    // the entered children are now a prefix of the four
    if (key[3].y != kNoEntry) st.push(stack_entry(key[3].x, key[3].y));
    if (key[2].y != kNoEntry) st.push(stack_entry(key[2].x, key[2].y));
    if (key[1].y != kNoEntry) st.push(stack_entry(key[1].x, key[1].y));
    return key[0].y != kNoEntry ? key[0].x : kInvalid;
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
#ifndef WR_SPEC_STEPS
#define WR_SPEC_STEPS 4
#endif
constexpr int kSpecSteps = WR_SPEC_STEPS;
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
            float tn[4];
            bool h[4];
            const uint4 ch = node_test(s, r, cur, cull_limit(ts.t), tn, h);
            if (COUNT) cnt.nodes++;
            uint2 key[4] = {make_uint2(ch.x, h[0] ? __float_as_uint(tn[0]) : kNoEntry),
                            make_uint2(ch.y, h[1] ? __float_as_uint(tn[1]) : kNoEntry),
                            make_uint2(ch.z, h[2] ? __float_as_uint(tn[2]) : kNoEntry),
                            make_uint2(ch.w, h[3] ? __float_as_uint(tn[3]) : kNoEntry)};
            const uint32_t next = sort_push(key, st);
            if (next != kInvalid) ts.cur = next;
            else pop = true;
        }
        if (pop) {
            more = false;
            const float lim = cull_limit(ts.t);
            while (st.sp > 0) {
                st.sp--;
                const uint64_t e = st.top();
                if (__uint_as_float(uint32_t(e >> 32)) <= lim) {
                    ts.cur = uint32_t(e);
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
            float tn[4];
            bool h[4];
            const uint4 ch = node_test(s, r, cur, cull_limit(ts.t), tn, h);
            if (COUNT) cnt.nodes++;
            const uint32_t c[4] = {ch.x, ch.y, ch.z, ch.w};
            uint32_t next = kInvalid;
#pragma unroll
            for (int i = 3; i >= 0; i--) {
                if (h[i]) {
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

// Whole traversal of one ray in speculative rounds (trav_round / trav_round_any):
// ties in t go to the lower leaf-order index.  The rounds' ballots count the lanes that are active, so lanes of a
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
