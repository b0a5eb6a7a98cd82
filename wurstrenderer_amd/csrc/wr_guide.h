// wr_guide.h -- path-guiding network on the device: one-blob encoding, the
// von Mises-Fisher (normalised spherical-Gaussian) mixture, and the wave-level
// BF16 MFMA forward pass.
//
// No reference counterpart: the reference has no learned guiding.  Its only SG
// convention is SgLobe::eval = a exp((1 - w.mu) lambda)
// (src/sphericalrepresentation/sphericalgaussian.h:10-21), which peaks at
// w = -mu; this file uses the normalised lobe
//   c(w) = lambda / (2 pi (1 - exp(-2 lambda))) exp(lambda (mu.w - 1)).
//
// Network: 32 inputs -> 64 -> 64 -> 64 -> 64 (ReLU) -> 40 raw outputs,
// 8 lobes x (weight logit, axis xyz, sharpness).  Mixture weights are
// softmax(logits), axes are normalised, sharpness = softplus(raw) clamped to
// [kLambdaMin, kLambdaMax].  Weights and activations are BF16, products
// accumulate in FP32 (v_mfma_f32_32x32x16_bf16), master weights are FP32.
//
// FP8 inference variant: weights and activations in e4m3 with FP32
// accumulation (v_mfma_f32_32x32x16_fp8_fp8), per-layer power-of-two weight
// scales and, per layer, a power-of-two activation scale for each record (from
// that record's own largest input), so a record's outputs never depend on the
// other records of its wave.  Scaled values stay within (64, 128] at the
// maximum, inside the normal range of e4m3.
//
// Register-resident layout.  The network runs transposed, Y^T = W^T X^T, so a
// wave's 64 records are the MFMA columns.  A layer's output D tile (lane
// (p, h) = (l % 32, l / 32) holds rows drow(v, h) = 8 (v / 4) + 4 h + v % 4 of
// column p in register v) feeds the next layer's B operand without leaving
// registers: B element e of k-step ks in lane half h is taken to be input
// feature kperm(ks, h, e) = 16 ks + 8 (e / 4) + 4 h + e % 4, which is exactly
// where the previous layer left it, and the weights are packed with the same
// permutation of the contraction index.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include "wr_device.h"

namespace wr {
namespace guide {

constexpr int kIn = 32, kHid = 64, kOut = 40, kLayers = 5, kLobes = 8;
constexpr float kLambdaMin = 0.05f, kLambdaMax = 2000.f;

// Layer l: weight W_l [K_l][N_l] (input-major), then bias b_l [N_l]; the FP32
// parameter vector stores W0, b0, W1, b1, ...
__host__ __device__ constexpr int layer_in(int l) { return l == 0 ? kIn : kHid; }
__host__ __device__ constexpr int layer_out(int l) { return l == kLayers - 1 ? kOut : kHid; }
// (closed forms: no recursion, which a runtime l would turn into a device call)
__host__ __device__ constexpr int w_offset(int l) {
    return l == 0 ? 0
                  : (kIn * kHid + kHid) + (l - 1) * (kHid * kHid + kHid) -
                        (l == kLayers ? (kHid * kHid + kHid) - (kHid * kOut + kOut) : 0);
}
__host__ __device__ constexpr int b_offset(int l) { return w_offset(l) + layer_in(l) * layer_out(l); }
constexpr int kNumParams = w_offset(kLayers);
static_assert(kNumParams == 17192, "parameter count");
static_assert(w_offset(1) == 32 * 64 + 64 && w_offset(4) == w_offset(3) + 64 * 64 + 64, "layer offsets");

// Packed BF16 operands, 64 lanes x 8 values per (tile, k-step):
//   forward  F_l[mt][ks]: lane (r, h), e -> W_l[kperm(ks, h, e)][32 mt + r]  (0 past N_l)
//            mt < 2 (outputs padded to 64), ks < K_l / 16
//   backward B_l[mt][ks]: lane (r, h), e -> W_l[32 mt + r][kperm(ks, h, e)]  (0 past N_l)
//            mt < K_l / 32, ks < 4 (outputs padded to 64); layers 1..4
__host__ __device__ constexpr int kperm(int ks, int h, int e) { return 16 * ks + 8 * (e >> 2) + 4 * h + (e & 3); }
__host__ __device__ constexpr int fwd_offset(int l) { return l == 0 ? 0 : 2 * (kIn / 16) * 64 + (l - 1) * 2 * (kHid / 16) * 64; }
constexpr int kFwdOps = fwd_offset(kLayers);
__host__ __device__ constexpr int bwd_offset(int l) { return l <= 1 ? 0 : (l - 1) * (kHid / 32) * 4 * 64; }
constexpr int kBwdOps = bwd_offset(kLayers);
// Row of D register v in lane half h.
__host__ __device__ constexpr int drow(int v, int h) { return 8 * (v >> 2) + 4 * h + (v & 3); }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
// An MFMA operand: 8 BF16 values packed two per 32-bit word (element e in word
// e / 2, low half first), so each operand is exactly 4 registers.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
// An FP8 operand: the same 8 elements as a BF16 operand, one e4m3 byte each
// (element e in byte e), 2 registers.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ f32x16 mfma(u32x4 a, u32x4 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}
__device__ __forceinline__ __bf16 bf(float x) { return __float2bfloat16(x); }
__device__ __forceinline__ float fb(__bf16 x) { return __bfloat162float(x); }
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
// Two FP32 -> two BF16 (round to nearest even) in one word, lo in the low half:
// one v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint32_t pack2(float lo, float hi) {
    const f32x2 v = {lo, hi};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}
typedef short s16x2 __attribute__((ext_vector_type(2)));
// pack2 of max(lo, 0) and max(hi, 0): a negative BF16 (or -0) has its sign bit
// set and is negative as an int16, so one v_pk_max_i16 against 0 applies the
// ReLU to both halves after the conversion.
__device__ __forceinline__ uint32_t pack2_relu(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2, pack2(lo, hi)), s16x2{0, 0}));
}
// c + both BF16 halves of w: one v_dot2_f32_bf16 against (1, 1)
__device__ __forceinline__ float add_halves(uint32_t w, float c) {
    return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, w), __builtin_bit_cast(bf16x2, 0x3F803F80u), c, false);
}

// Device view of a network.
// Biases of all layers packed at 64 per layer (the last has 40).
constexpr int kBiasPacked = 64 * (kLayers - 1) + kOut;

struct Net {
    const u32x4 *fwd;     // kFwdOps packed forward operands
    const u32x4 *bwd;     // kBwdOps packed backward operands
    const float *bias;    // kBiasPacked FP32 biases, layer l at 64 l
    const u32x2 *fwd8;    // kFwdOps packed e4m3 forward operands (weights / wscale[l])
    const float *wscale;  // kLayers power-of-two weight scales
    float box[6];         // position encoding: scene bounds lo xyz, then 1 / extent xyz
};

// Training records written by the renderer, one per recorded path vertex.
struct Records {
    float4 *pos;       // xyz: vertex position; w: sample index in the pass (bits)
    float4 *nrm;       // xyz: normal facing the incoming ray
    float4 *wo;        // xyz: direction to the previous vertex
    float4 *wi;        // xyz: sampled direction; w: training weight (set after the pass)
    float4 *thr;       // xyz: path throughput before the vertex
    float4 *rad;       // xyz: path radiance when the next vertex is shaded
    uint32_t *count;   // records appended (may exceed capacity; the excess is dropped)
    uint32_t capacity;
};

// What the inference kernel leaves for the shading kernel, per entry of the
// path state: the direction it drew (from the network with probability alpha,
// else from the BSDF or phase function) with the network's density there, and
// the network's density in the direction of the vertex's light sample.
struct Draw {
    float4 *dir;   // xyz: sampled direction, w: q(direction)
    float *q_nee;  // q(direction to the light sample)
};

// What the shading kernel needs for guided sampling.
struct Shade {
    Draw draw;
    Records rec;
    float alpha;        // one-sample MIS probability of sampling the network (0: BSDF only)
    float record_prob;  // fraction of paths whose vertices are recorded
    int collect;        // record training data
    uint32_t max_bounce;  // vertices at bounce >= this are neither guided nor recorded
};

// Hardware exp2 / log2 (v_exp_f32 / v_log_f32, about 1 ulp) for the network's
// encoding and mixture: the features are rounded to BF16 anyway, and the
// mixture's pdf and sampler read the same decoded values.
__device__ __forceinline__ float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
__device__ __forceinline__ float flog(float x) { return __builtin_amdgcn_logf(x) * 0.69314718055994531f; }

// ---------------------------------------------------------------- encoding
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
// Octahedral map of a unit vector to [0,1]^2 (Engelhardt and Dachsbacher 2008),
// y up: project onto |x| + |y| + |z| = 1, fold the lower half over the diagonals.
__device__ __forceinline__ void oct01(float3 d, float &a, float &b) {
    const float r = frcp(fabsf(d.x) + fabsf(d.y) + fabsf(d.z));
    float u = d.x * r, v = d.z * r;
    if (d.y < 0.f) {
        const float uu = (1.f - fabsf(v)) * (u >= 0.f ? 1.f : -1.f);
        v = (1.f - fabsf(u)) * (v >= 0.f ? 1.f : -1.f);
        u = uu;
    }
    a = fminf(fmaxf(0.5f * u + 0.5f, 0.f), 1.f);
    b = fminf(fmaxf(0.5f * v + 0.5f, 0.f), 1.f);
}
// The 7 encoded scalars of a record, all in [0,1]: position in the scene box
// (3), normal (2), direction to the previous vertex (2).  box: lo xyz, then
// 1 / extent xyz.
struct Enc { float v[7]; };
__device__ __forceinline__ Enc encode(float3 pos, float3 n, float3 wo, const float box[6]) {
    Enc e;
    e.v[0] = fminf(fmaxf((pos.x - box[0]) * box[3], 0.f), 1.f);
    e.v[1] = fminf(fmaxf((pos.y - box[1]) * box[4], 0.f), 1.f);
    e.v[2] = fminf(fmaxf((pos.z - box[2]) * box[5], 0.f), 1.f);
    oct01(n, e.v[3], e.v[4]);
    oct01(wo, e.v[5], e.v[6]);
    return e;
}
// One-blob feature f of 32 (Mueller et al. 2019, "Neural Importance
// Sampling"): a Gaussian kernel one bin wide over k bins; position 3 x 6 bins,
// normal 2 x 4, direction 2 x 3.
__host__ __device__ constexpr int blob_dim(int f) { return f < 6 ? 0 : f < 12 ? 1 : f < 18 ? 2 : f < 22 ? 3 : f < 26 ? 4 : f < 29 ? 5 : 6; }
__host__ __device__ constexpr int blob_bins(int f) { return f < 18 ? 6 : f < 26 ? 4 : 3; }
__host__ __device__ constexpr float blob_centre(int f) {
    return (float(f < 18 ? f % 6 : f < 26 ? (f - 18) % 4 : (f - 26) % 3) + 0.5f) / float(blob_bins(f));
}
// Feature f0 (lane half h = 0) or f0 + 4 (h = 1) of the record, f0 a
// compile-time constant (the features of lane half h are kperm(ks, h, e) =
// kperm(ks, 0, e) + 4 h), so only three selects depend on h.
__device__ __forceinline__ float feature_h(const Enc &e, int f0, int h) {
    const int f1 = f0 + 4;
    const float x = h ? e.v[blob_dim(f1)] : e.v[blob_dim(f0)];
    const float c = h ? blob_centre(f1) : blob_centre(f0);
    const float k = h ? float(blob_bins(f1)) : float(blob_bins(f0));
    const float d = (x - c) * k;
    return fexp(-0.5f * d * d);
}

// ---------------------------------------------------------------- mixture
// Decoded lobes live in a row of kStageStride floats (LDS or registers):
// lobe k at 5k: (log w_k + log norm_k, mu_k xyz, lambda_k), with
// norm = lambda / (2 pi (1 - exp(-2 lambda))).
constexpr int kStageStride = 41;  // odd: the 64 rows of a wave fall in distinct banks
// softplus and log norm(lambda) from the hardware exp2 / log2 (lambda >= kLambdaMin,
// so 1 - e^{-2 lambda} >= 0.095 and the difference loses no digits that matter)
__device__ __forceinline__ float softplus_fast(float x) { return x > 15.f ? x : flog(1.f + fexp(x)); }
__device__ __forceinline__ float log_vmf_norm_fast(float lam) {
    return flog(lam * (1.f / (2.f * kPi)) * frcp(1.f - fexp(-2.f * lam)));
}

// Raw outputs -> decoded lobes, in place; w (if given) receives the mixture
// weights, which the sampler picks lobes by.
__device__ __forceinline__ void decode(float *o, float *w = nullptr) {
    float mx = -INFINITY;
#pragma unroll
    for (int k = 0; k < kLobes; k++) mx = fmaxf(mx, o[5 * k]);
    float e[kLobes], s = 0.f;
#pragma unroll
    for (int k = 0; k < kLobes; k++) {
        e[k] = fexp(o[5 * k] - mx);
        s += e[k];
    }
    const float lse = mx + flog(s), rs = frcp(s);
#pragma unroll
    for (int k = 0; k < kLobes; k++) {
        if (w) w[k] = e[k] * rs;
        const float3 a = f3(o[5 * k + 1], o[5 * k + 2], o[5 * k + 3]);
        const float n2 = dot3(a, a);
        const float3 mu = n2 > 1e-24f ? a * __builtin_amdgcn_rsqf(n2) : f3(0.f, 1.f, 0.f);
        const float lam = fminf(fmaxf(softplus_fast(o[5 * k + 4]), kLambdaMin), kLambdaMax);
        o[5 * k] = o[5 * k] - lse + log_vmf_norm_fast(lam);
        o[5 * k + 1] = mu.x;
        o[5 * k + 2] = mu.y;
        o[5 * k + 3] = mu.z;
        o[5 * k + 4] = lam;
    }
}
__device__ __forceinline__ float pdf(const float *L, float3 w) {
    float p = 0.f;
#pragma unroll
    for (int k = 0; k < kLobes; k++)
        p += fexp(L[5 * k] + L[5 * k + 4] * (L[5 * k + 1] * w.x + L[5 * k + 2] * w.y + L[5 * k + 3] * w.z - 1.f));
    return p;
}
// Lobe by weight w (u0), then the vMF around its axis (Wood 1994, on the
// 2-sphere): cos = 1 + log(1 - u1 + u1 exp(-2 lambda)) / lambda, azimuth 2 pi u2.
// The lobe is picked with selects, so L and w may live in registers.
__device__ __forceinline__ float3 sample(const float *L, const float *w, float u0, float u1, float u2) {
    float3 mu = f3(L[5 * (kLobes - 1) + 1], L[5 * (kLobes - 1) + 2], L[5 * (kLobes - 1) + 3]);
    float lam = L[5 * (kLobes - 1) + 4];
    float c = 0.f;
    bool found = false;
#pragma unroll
    for (int k = 0; k < kLobes - 1; k++) {
        c += w[k];
        const bool take = !found && u0 < c;
        mu = take ? f3(L[5 * k + 1], L[5 * k + 2], L[5 * k + 3]) : mu;
        lam = take ? L[5 * k + 4] : lam;
        found = found || take;
    }
    const float ct = fminf(fmaxf(1.f + flog(1.f - u1 + u1 * fexp(-2.f * lam)) * frcp(lam), -1.f), 1.f);
    const float st = sqrtf(fmaxf(0.f, 1.f - ct * ct));
    float s, co;
    sincospif(2.f * u2, &s, &co);
    return normalize3(to_world(make_frame(mu), f3(st * co, ct, st * s)));
}

// ---------------------------------------------------------------- forward pass
// Input operands of layer 0 for the wave's 64 records: record r = 32 nt + p is
// encoded in lane r; lane (p, h) builds its 16 features of records p and 32 + p.
__device__ __forceinline__ void input_operands(const Enc &mine, u32x4 x[2][2] /* [ks][nt] */) {
    const int lane = __lane_id(), p = lane & 31, h = lane >> 5;
#pragma unroll
    for (int nt = 0; nt < 2; nt++) {
        Enc e;
#pragma unroll
        for (int j = 0; j < 7; j++) e.v[j] = __shfl(mine.v[j], 32 * nt + p);
#pragma unroll
        for (int ks = 0; ks < 2; ks++)
#pragma unroll
            for (int i = 0; i < 4; i++)
                x[ks][nt][i] = pack2(feature_h(e, kperm(ks, 0, 2 * i), h), feature_h(e, kperm(ks, 0, 2 * i + 1), h));
    }
}

// Output tile mt (32 outputs) of a layer for the wave's 64 records:
// acc[nt] = b + (W_l^T X^T)[32 mt .. 32 mt + 31][32 nt .. 32 nt + 31]; the
// accumulators start at the bias (N: outputs of the layer).
template <int K>
__device__ __forceinline__ void tile_fwd(const u32x4 *ops, const float *bias, int N, const u32x4 (*x)[2], int mt,
                                         f32x16 acc[2]) {
    const int lane = __lane_id(), h = lane >> 5;
    const float *bb = bias + 4 * h;  // per-lane base, compile-time offsets
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int o0 = 32 * mt + drow(v, 0);
        acc[0][v] = o0 + 4 * h < N ? bb[o0] : 0.f;
    }
    acc[1] = acc[0];
#pragma unroll
    for (int ks = 0; ks < K / 16; ks++) {
        const u32x4 a = ops[(mt * (K / 16) + ks) * 64 + lane];
        acc[0] = mfma(a, x[ks][0], acc[0]);
        acc[1] = mfma(a, x[ks][1], acc[1]);
    }
}

// ReLU of output tile mt into operands 2 mt, 2 mt + 1 of the next layer:
// xn[ks][nt] element e = relu(acc[nt][8 (ks % 2) + e]).
__device__ __forceinline__ void tile_to_operands(const f32x16 acc[2], int mt, u32x4 xn[4][2]) {
#pragma unroll
    for (int kk = 0; kk < 2; kk++)
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
                xn[2 * mt + kk][nt][i] = pack2_relu(acc[nt][8 * kk + 2 * i], acc[nt][8 * kk + 2 * i + 1]);
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Output row pair of one record: lane (p, h) holds, in r0 and r1, the rows
// 32 mt + drow(v, h) of records p and 32 + p of output tile mt.
// v_permlane32_swap exchanges the upper half-wave of r0 with the lower
// half-wave of r1, after which lane (p, h) -- record 32 h + p -- holds rows
// 32 mt + drow(v, 0) (lo) and 32 mt + drow(v, 0) + 4 (hi) of its own record.
__device__ __forceinline__ void rows_to_record(float r0, float r1, float &lo, float &hi) {
    const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(r0), __float_as_uint(r1), false, false);
    lo = __uint_as_float(sw[0]);
    hi = __uint_as_float(sw[1]);
}

// Forward pass of the wave's 64 records (record r is the one encoded in lane
// r).  Lane r receives the raw outputs of record r in out[0 .. kOut).  Must be
// called by all 64 lanes of the wave.
__device__ __forceinline__ void forward(const Net &net, const Enc &mine, float out[kOut]) {
    u32x4 x[4][2];
    {
        u32x4 x0[2][2];
        input_operands(mine, x0);
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            f32x16 acc[2];
            tile_fwd<kIn>(net.fwd + fwd_offset(0), net.bias, kHid, x0, mt, acc);
            tile_to_operands(acc, mt, x);
        }
    }
    // hidden layers 1..3 (same shape): rolled, so one layer's operands are live at a time
#pragma unroll 1
    for (int l = 1; l < kLayers - 1; l++) {
        u32x4 xn[4][2];
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            f32x16 acc[2];
            tile_fwd<kHid>(net.fwd + fwd_offset(l), net.bias + 64 * l, kHid, x, mt, acc);
            tile_to_operands(acc, mt, xn);
        }
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++) x[ks][nt] = xn[ks][nt];
    }
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
        f32x16 acc[2];
        tile_fwd<kHid>(net.fwd + fwd_offset(kLayers - 1), net.bias + 64 * (kLayers - 1), kOut, x, mt, acc);
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int o0 = 32 * mt + drow(v, 0);
            if (o0 < kOut) {
                float lo, hi;
                rows_to_record(acc[0][v], acc[1][v], lo, hi);
                out[o0] = lo;
                if (o0 + 4 < kOut) out[o0 + 4] = hi;
            }
        }
    }
}

// ---------------------------------------------------------------- FP8 forward pass
__device__ __forceinline__ f32x16 mfma8(u32x2 a, u32x2 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(__builtin_bit_cast(long, a), __builtin_bit_cast(long, b), c, 0, 0,
                                                      0);
}
// Four FP32 values (|x| <= 128) -> four e4m3 bytes, round to nearest even.
__device__ __forceinline__ uint32_t pack4_fp8(float a, float b, float c, float d) {
    return uint32_t(__builtin_amdgcn_cvt_pk_fp8_f32(c, d, __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false), true));
}
// Power-of-two scale of a set with largest magnitude m: k = ceil(log2 m) - 7
// (from the FP32 bits, exactly), so m 2^-k lies in (64, 128]; k = 0 for m = 0.
// Returns 2^k; the quantiser multiplies by 2^-k (exact).
__device__ __forceinline__ int pow2_exponent(float m) {
    const uint32_t u = __float_as_uint(m);
    if (!(m > 0.f)) return 0;
    const int e = int(u >> 23) - 127 + ((u & 0x7FFFFFu) ? 1 : 0);
    const int k = e - 7;
    return k < -100 ? -100 : (k > 100 ? 100 : k);
}
__device__ __forceinline__ float pow2f(int k) { return __uint_as_float(uint32_t(127 + k) << 23); }

// Quantises 8 values (the elements of one operand) scaled by inv.
__device__ __forceinline__ u32x2 quant8(const float *v, float inv) {
    u32x2 r;
    r[0] = pack4_fp8(v[0] * inv, v[1] * inv, v[2] * inv, v[3] * inv);
    r[1] = pack4_fp8(v[4] * inv, v[5] * inv, v[6] * inv, v[7] * inv);
    return r;
}

// One layer on e4m3 operands: acc[mt][nt] = W8 X8 over K inputs, both output tiles.
template <int K>
__device__ __forceinline__ void layer8(const u32x2 *ops, const u32x2 (*x)[2], f32x16 acc[2][2]) {
    const int lane = __lane_id();
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[mt][nt][v] = 0.f;
#pragma unroll
        for (int ks = 0; ks < K / 16; ks++) {
            const u32x2 a = ops[(mt * (K / 16) + ks) * 64 + lane];
            acc[mt][0] = mfma8(a, x[ks][0], acc[mt][0]);
            acc[mt][1] = mfma8(a, x[ks][1], acc[mt][1]);
        }
    }
}

// Dequantised pre-activation of output tile mt: acc * (wscale * xscale[nt]) + bias.
// Hidden layers then take relu, find each record's largest input to the next
// layer (its 64 values sit in the lane's two tiles and in lane p ^ 32) and
// quantise with that record's scale.
__device__ __forceinline__ void hidden_to_operands8(f32x16 acc[2][2], const float *bias, float ws, const float xs[2],
                                                    u32x2 xn[4][2], float xsn[2]) {
    const int lane = __lane_id(), h = lane >> 5;
    const float *bb = bias + 4 * h;
#pragma unroll
    for (int nt = 0; nt < 2; nt++) {
        const float sc = ws * xs[nt];
        float m = 0.f;
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const float y = fmaxf(fmaf(acc[mt][nt][v], sc, bb[32 * mt + drow(v, 0)]), 0.f);
                acc[mt][nt][v] = y;
                m = fmaxf(m, y);
            }
        m = fmaxf(m, __shfl_xor(m, 32));
        const int k = pow2_exponent(m);
        const float inv = pow2f(-k);
        xsn[nt] = pow2f(k);
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int kk = 0; kk < 2; kk++) {
                float v8[8];
#pragma unroll
                for (int e = 0; e < 8; e++) v8[e] = acc[mt][nt][8 * kk + e];
                xn[2 * mt + kk][nt] = quant8(v8, inv);
            }
    }
}

// forward() on e4m3 weights and activations: same inputs and outputs.
__device__ __forceinline__ void forward8(const Net &net, const Enc &mine, float out[kOut]) {
    const int lane = __lane_id(), p = lane & 31, h = lane >> 5;
    u32x2 x[4][2];
    float xs[2];
    {
        // layer-0 inputs: the one-blob features of records p and 32 + p
#pragma unroll
        for (int nt = 0; nt < 2; nt++) {
            Enc e;
#pragma unroll
            for (int j = 0; j < 7; j++) e.v[j] = __shfl(mine.v[j], 32 * nt + p);
            float f[16];
            float m = 0.f;
#pragma unroll
            for (int ks = 0; ks < 2; ks++)
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    f[8 * ks + i] = feature_h(e, kperm(ks, 0, i), h);
                    m = fmaxf(m, f[8 * ks + i]);
                }
            m = fmaxf(m, __shfl_xor(m, 32));
            const int k = pow2_exponent(m);
            xs[nt] = pow2f(k);
            const float inv = pow2f(-k);
            x[0][nt] = quant8(f, inv);
            x[1][nt] = quant8(f + 8, inv);
        }
        f32x16 acc[2][2];
        layer8<kIn>(net.fwd8 + fwd_offset(0), x, acc);
        float xsn[2];
        u32x2 xn[4][2];
        hidden_to_operands8(acc, net.bias, net.wscale[0], xs, xn, xsn);
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++) x[ks][nt] = xn[ks][nt];
        xs[0] = xsn[0];
        xs[1] = xsn[1];
    }
#pragma unroll 1
    for (int l = 1; l < kLayers - 1; l++) {
        f32x16 acc[2][2];
        layer8<kHid>(net.fwd8 + fwd_offset(l), x, acc);
        float xsn[2];
        u32x2 xn[4][2];
        hidden_to_operands8(acc, net.bias + 64 * l, net.wscale[l], xs, xn, xsn);
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++) x[ks][nt] = xn[ks][nt];
        xs[0] = xsn[0];
        xs[1] = xsn[1];
    }
    f32x16 acc[2][2];
    layer8<kHid>(net.fwd8 + fwd_offset(kLayers - 1), x, acc);
    const float ws = net.wscale[kLayers - 1];
    const float *bb = net.bias + 64 * (kLayers - 1) + 4 * h;
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int o0 = 32 * mt + drow(v, 0);
            if (o0 < kOut) {
                // this lane's row is o0 + 4 h (its bias bb[o0]); rows past kOut are padding
                const float b = o0 + 4 * h < kOut ? bb[o0] : 0.f;
                float lo, hi;
                rows_to_record(fmaf(acc[mt][0][v], ws * xs[0], b), fmaf(acc[mt][1][v], ws * xs[1], b), lo, hi);
                out[o0] = lo;
                if (o0 + 4 < kOut) out[o0 + 4] = hi;
            }
        }
}

// ---------------------------------------------------------------- loss gradient
// One record: loss = -w log q(wi).  Writes dloss/dout[40] to g (which may be
// o itself) and returns the loss.  Computed in log space: lobes underflow far
// from their axis.
__device__ __forceinline__ float loss_grad(const float *o, float3 wi, float w, float *g) {
    // hardware exp2 / log2 / reciprocal throughout (about 1 ulp each): the
    // outputs come from a BF16 network, so these errors are far below its own
    float mx = -INFINITY;
    for (int k = 0; k < kLobes; k++) mx = fmaxf(mx, o[5 * k]);
    float s = 0.f;
    for (int k = 0; k < kLobes; k++) s += fexp(o[5 * k] - mx);
    const float lse = mx + flog(s);
    float lc[kLobes], lam[kLobes], inv[kLobes], dl[kLobes], cs[kLobes];
    float3 mu[kLobes];
    float lq_max = -INFINITY;
    for (int k = 0; k < kLobes; k++) {
        const float3 a = f3(o[5 * k + 1], o[5 * k + 2], o[5 * k + 3]);
        const float n2 = dot3(a, a);
        inv[k] = n2 > 1e-24f ? __builtin_amdgcn_rsqf(n2) : 0.f;  // 1 / |a|, 0 for a degenerate axis
        mu[k] = n2 > 1e-24f ? a * inv[k] : f3(0.f, 1.f, 0.f);
        const float raw = o[5 * k + 4];
        const float sp = softplus_fast(raw);
        lam[k] = fminf(fmaxf(sp, kLambdaMin), kLambdaMax);
        dl[k] = (sp > kLambdaMin && sp < kLambdaMax) ? frcp(1.f + fexp(-raw)) : 0.f;  // dlambda/draw
        cs[k] = dot3(mu[k], wi);
        lc[k] = (o[5 * k] - lse) + log_vmf_norm_fast(lam[k]) + lam[k] * (cs[k] - 1.f);
        lq_max = fmaxf(lq_max, lc[k]);
    }
    float sq = 0.f;
    for (int k = 0; k < kLobes; k++) sq += fexp(lc[k] - lq_max);
    const float logq = lq_max + flog(sq);
    for (int k = 0; k < kLobes; k++) {
        const float gam = fexp(lc[k] - logq);  // responsibility of lobe k
        g[5 * k] = -w * (gam - fexp(o[5 * k] - lse));
        // d log c / d lambda = 1/lambda - 2 / (exp(2 lambda) - 1) + (mu.wi - 1)
        const float dlogc = frcp(lam[k]) - 2.f * frcp(fexp(2.f * lam[k]) - 1.f) + (cs[k] - 1.f);
        g[5 * k + 4] = -w * gam * dlogc * dl[k];
        // d log c / d mu = lambda wi, through mu = a / |a| (zero for a degenerate axis)
        const float3 gm = wi * (gam * lam[k]);
        const float3 ga = (gm - mu[k] * dot3(mu[k], gm)) * inv[k];
        g[5 * k + 1] = -w * ga.x;
        g[5 * k + 2] = -w * ga.y;
        g[5 * k + 3] = -w * ga.z;
    }
    return -w * logq;
}

}  // namespace guide
}  // namespace wr
