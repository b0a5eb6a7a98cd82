// wr_guide.hip -- training and test kernels of the path-guiding network, and
// its C-ABI (no reference counterpart; the reference has no guiding).
//
//   pack      FP32 master weights -> BF16 MFMA operands (forward and backward)
//   train     per wave of 64 records: forward (MFMA), loss gradient, backward
//             data (MFMA) and weight gradients (MFMA over the records), summed
//             per block in LDS and written as one partial per block
//   adam      sums the block partials and applies Adam to the FP32 master
//   finalize  training weight of each record from its path's final radiance
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "wr_guide.h"
#include "wr_host.h"

using namespace wr;
using namespace wr::guide;

namespace {

constexpr int kXt = 72;                  // LDS row stride (bf16) of the transposed tiles
constexpr int kMaxTrainBlocks = 256;
constexpr uint32_t kTestBlock = 64;

// ---------------------------------------------------------------- pack
// `sc` scales every parameter as it is read (1, or the debiasing factor of the
// parameter average); a null `bwd` skips the backward operands.
__device__ __forceinline__ void pack_bf16(int j, const float *params, float sc, u32x4 *fwd, u32x4 *bwd, float *bias) {
    if (j < kBiasPacked) bias[j] = sc * params[b_offset(j / 64) + j % 64];
    if (j < kFwdOps) {
        int l = 0;
        while (l + 1 < kLayers && j >= fwd_offset(l + 1)) l++;
        const int K = layer_in(l), N = layer_out(l), nks = K / 16;
        const int jj = j - fwd_offset(l);
        const int mt = jj / (nks * 64), ks = (jj / 64) % nks, lane = jj % 64;
        const int o = 32 * mt + (lane & 31), h = lane >> 5;
        u32x4 v;
        for (int i = 0; i < 4; i++)
            v[i] = o < N ? pack2(sc * params[w_offset(l) + kperm(ks, h, 2 * i) * N + o],
                                 sc * params[w_offset(l) + kperm(ks, h, 2 * i + 1) * N + o])
                         : 0u;
        fwd[j] = v;
    } else if (bwd && j < kFwdOps + kBwdOps) {
        const int jb = j - kFwdOps;
        int l = 1;
        while (l + 1 < kLayers && jb >= bwd_offset(l + 1)) l++;
        const int N = layer_out(l);
        const int jj = jb - bwd_offset(l);
        const int mt = jj / (4 * 64), ks = (jj / 64) % 4, lane = jj % 64;
        const int i = 32 * mt + (lane & 31), h = lane >> 5;
        u32x4 v;
        for (int k = 0; k < 4; k++) {
            const int o0 = kperm(ks, h, 2 * k), o1 = kperm(ks, h, 2 * k + 1);
            v[k] = pack2(o0 < N ? sc * params[w_offset(l) + i * N + o0] : 0.f,
                         o1 < N ? sc * params[w_offset(l) + i * N + o1] : 0.f);
        }
        bwd[jb] = v;
    }
}

// FP32 master weights -> e4m3 forward operands, one block per layer: the
// layer's largest |weight| sets its power-of-two scale (wscale[l] = 2^k, the
// operands hold W / 2^k, at most 128 in magnitude).
constexpr int kPack8Block = 256;
__device__ __forceinline__ void pack_e4m3(int l, float *red, const float *params, float sc, u32x2 *fwd8, float *wscale) {
    const int K = layer_in(l), N = layer_out(l), nks = K / 16;
    const float *W = params + w_offset(l);
    float m = 0.f;
    for (int j = threadIdx.x; j < K * N; j += kPack8Block) m = fmaxf(m, fabsf(sc * W[j]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kPack8Block / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const int k = pow2_exponent(red[0]);
    const float inv = pow2f(-k);
    if (threadIdx.x == 0) wscale[l] = pow2f(k);
    const int n_ops = fwd_offset(l + 1) - fwd_offset(l);
    for (int jj = threadIdx.x; jj < n_ops; jj += kPack8Block) {
        const int mt = jj / (nks * 64), ks = (jj / 64) % nks, lane = jj % 64;
        const int o = 32 * mt + (lane & 31), h = lane >> 5;
        float v[8];
        for (int e = 0; e < 8; e++) v[e] = o < N ? sc * W[kperm(ks, h, e) * N + o] : 0.f;
        fwd8[fwd_offset(l) + jj] = quant8(v, inv);
    }
}
// Both packings in one launch: blocks 0 .. kLayers-1 quantise one layer each,
// the rest convert 256 BF16 operands each.
__global__ __launch_bounds__(kPack8Block) void k_pack(const float *params, float sc, u32x4 *fwd, u32x4 *bwd, float *bias,
                                                    u32x2 *fwd8, float *wscale) {
    __shared__ float red[kPack8Block];
    if (blockIdx.x < kLayers) pack_e4m3(blockIdx.x, red, params, sc, fwd8, wscale);
    else pack_bf16((blockIdx.x - kLayers) * kPack8Block + threadIdx.x, params, sc, fwd, bwd, bias);
}

__global__ void k_fill_u32_guide(uint32_t *p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}
__global__ void k_fill_f64_guide(double *p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0;
}

// ---------------------------------------------------------------- train
struct TrainArgs {
    const float4 *pos, *nrm, *wo, *wi;  // records; wi.w = weight
    uint32_t n;                         // records available
    uint32_t start, batch;              // this step uses records (start + j) % n, j < batch
    float wscale;                       // loss = sum_j wscale * w_j * (-log q_j)
    float *partials;                    // [gridDim.x][kNumParams]
    float *loss;                        // [gridDim.x]
};

// Four waves per block (one per SIMD of a CU) step through the records in
// lockstep, a tile of 64 per wave.  A wave keeps every layer's input operands
// in registers (it has a SIMD's whole register file), so LDS holds the current
// layer's transposed tiles (which the loss stage aliases), 4 x 18 KiB, and a
// copy of the packed weights and biases, 69 KiB: at one wave per SIMD nothing
// hides a load's latency, and LDS answers several times sooner than L2.  Each
// 32 x 32 tile of a weight gradient belongs to one wave, which reduces it over
// all four waves' staged tiles with MFMA and adds it into the block's slice of
// the partial sums in global memory: one writer per address, no atomics.
constexpr int kTrainWaves = 4;
struct TrainLds {
    union {
        struct {
            uint16_t xt[64 * kXt];   // X_l^T of the current layer (BF16 bits): feature-major, records contiguous
            uint16_t dyt[64 * kXt];  // dY_l^T, likewise
        } t;
        float stage[64 * kStageStride];  // raw outputs, then d loss / d outputs
    };
};

// X^T of one layer's input operands into LDS (record-contiguous rows).
// (Addresses are a per-lane base plus compile-time offsets, so they fold into
// the LDS instructions' immediates instead of occupying registers.)
template <int KS>
__device__ __forceinline__ void store_xt(uint16_t *xt, const u32x4 (*x)[2]) {
    const int lane = __lane_id(), p = lane & 31, h = lane >> 5;
    uint16_t *base = xt + 4 * h * kXt + p;
#pragma unroll
    for (int ks = 0; ks < KS; ks++)
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
#pragma unroll
            for (int e = 0; e < 8; e++)
                base[kperm(ks, 0, e) * kXt + 32 * nt] = uint16_t(x[ks][nt][e >> 1] >> (16 * (e & 1)));
}

// Backward through layer l of the block's four tiles (one per wave).  dY_l
// arrives as packed BF16 operands (dy[ks][nt] element e = dY of output
// kperm(ks, h, e) of record 32 nt + p, the values every product below uses);
// the layer's input operands are in x.  Adds dW_l and db_l into grad (stores
// them when first) and, for
// l > 0, leaves dY_{l-1} (the input's ReLU mask applied) packed in dy.  Every
// wave of the block calls it together (block barriers inside).
template <int l>
__device__ __forceinline__ void backward_layer(const Net &net, TrainLds *wl, float *grad, bool first, float (*bpart)[64],
                                               const u32x4 (*x)[2], u32x4 dy[4][2]) {
    constexpr int K = layer_in(l), N = layer_out(l);
    const int lane = __lane_id(), p = lane & 31, h = lane >> 5, wave = threadIdx.x >> 6;
    TrainLds &L = wl[wave];
    // weight gradient dW[i][o] = sum_r X[r][i] dY[r][o] over the block's 256
    // records: output tile (mt, nt) = (wave / 2, wave % 2) belongs to this wave.
    // Its running sum is the product's initial accumulator, loaded here so that
    // the staging and the barrier hide the load's latency.
    constexpr int kTiles = 2 * (K / 32);
    const int mt = wave >> 1, nt = wave & 1;
    const bool mine = wave < kTiles && 32 * nt + p < N;
    float *gb = grad + w_offset(l) + (32 * mt + 4 * h) * N + 32 * nt + p;
    f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; v++) acc[v] = !first && mine ? gb[drow(v, 0) * N] : 0.f;
    store_xt<4>(L.t.dyt, dy);
    store_xt<K / 16>(L.t.xt, x);
    wave_sync_lds();
    // this wave's share of the bias gradient: row sums of its dY^T
    if (lane < N) {
        float s[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int r = 0; r < 64; r += 8) {
            const u32x4 v = *reinterpret_cast<const u32x4 *>(L.t.dyt + lane * kXt + r);
#pragma unroll
            for (int i = 0; i < 4; i++) s[i] = add_halves(v[i], s[i]);
        }
        bpart[wave][lane] = (s[0] + s[1]) + (s[2] + s[3]);
    }
    __syncthreads();
    if (wave < kTiles) {
#pragma unroll
        for (int w2 = 0; w2 < kTrainWaves; w2++)
#pragma unroll
            for (int ks = 0; ks < 4; ks++) {
                const u32x4 av = *reinterpret_cast<const u32x4 *>(wl[w2].t.xt + (32 * mt + p) * kXt + 16 * ks + 8 * h);
                const u32x4 bv = *reinterpret_cast<const u32x4 *>(wl[w2].t.dyt + (32 * nt + p) * kXt + 16 * ks + 8 * h);
                acc = mfma(av, bv, acc);
            }
        if (mine) {
#pragma unroll
            for (int v = 0; v < 16; v++) gb[drow(v, 0) * N] = acc[v];
        }
    }
    // bias: the wave with no weight tile for layer 0, else one per layer in turn
    if (wave == (kTiles == 2 ? 2 : l & 3) && lane < N) {
        const float s = (bpart[0][lane] + bpart[1][lane]) + (bpart[2][lane] + bpart[3][lane]);
        grad[b_offset(l) + lane] = first ? s : grad[b_offset(l) + lane] + s;
    }
    if constexpr (l > 0) {
        // backward data: dX^T = W_l dY^T; the input's ReLU mask sits in the same lane
        // and slot as the result, which is packed into dy for the layer below
        const u32x4 *ops = net.bwd + bwd_offset(l);
        u32x4 nd[4][2];
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            f32x16 d0 = f32x16{}, d1 = f32x16{};
#pragma unroll
            for (int ks = 0; ks < 4; ks++) {
                const u32x4 av = ops[(mt * 4 + ks) * 64 + lane];
                d0 = mfma(av, dy[ks][0], d0);
                d1 = mfma(av, dy[ks][1], d1);
            }
#pragma unroll
            for (int v = 0; v < 16; v += 2) {
                // ReLU outputs are >= 0: the mask is "BF16 bits non-zero"
                const u32x4 &m0 = x[2 * mt + (v >> 3)][0], &m1 = x[2 * mt + (v >> 3)][1];
                const int w = (v & 7) >> 1;
                const float a0 = (m0[w] & 0xFFFFu) ? d0[v] : 0.f, b0 = (m0[w] >> 16) ? d0[v + 1] : 0.f;
                const float a1 = (m1[w] & 0xFFFFu) ? d1[v] : 0.f, b1 = (m1[w] >> 16) ? d1[v + 1] : 0.f;
                nd[2 * mt + (v >> 3)][0][w] = pack2(a0, b0);
                nd[2 * mt + (v >> 3)][1][w] = pack2(a1, b1);
            }
        }
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++) dy[ks][nt] = nd[ks][nt];
    }
    __syncthreads();  // every wave's tiles are read before the next layer (or record tile) overwrites them
}

// One hidden layer of the wave's 64 records: input operands x (K inputs) ->
// output operands xn (bias + ReLU, BF16).
template <int K>
__device__ __forceinline__ void layer_to(const Net &net, int l, const u32x4 (*x)[2], u32x4 xn[4][2]) {
#pragma unroll
    for (int mt = 0; mt < 2; mt++) {
        f32x16 acc[2];
        tile_fwd<K>(net.fwd + fwd_offset(l), net.bias + 64 * l, kHid, x, mt, acc);
        tile_to_operands(acc, mt, xn);
    }
}

__global__ __launch_bounds__(64 * kTrainWaves) void k_train(Net gnet, TrainArgs a) {
    __shared__ __attribute__((aligned(16))) TrainLds wl[kTrainWaves];
    __shared__ __attribute__((aligned(16))) u32x4 wfwd[kFwdOps];
    __shared__ __attribute__((aligned(16))) u32x4 wbwd[kBwdOps];
    __shared__ float wbias[kBiasPacked];
    __shared__ float bpart[kTrainWaves][64];
    __shared__ float wave_loss[kTrainWaves];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, p = lane & 31, h = lane >> 5;
    for (int j = threadIdx.x; j < kFwdOps; j += blockDim.x) wfwd[j] = gnet.fwd[j];
    for (int j = threadIdx.x; j < kBwdOps; j += blockDim.x) wbwd[j] = gnet.bwd[j];
    for (int j = threadIdx.x; j < kBiasPacked; j += blockDim.x) wbias[j] = gnet.bias[j];
    __syncthreads();
    Net net = gnet;
    net.fwd = wfwd;
    net.bwd = wbwd;
    net.bias = wbias;
    float *grad = a.partials + size_t(blockIdx.x) * kNumParams;
    TrainLds &L = wl[wave];
    float loss = 0.f;
    const uint32_t tiles = (a.batch + 63) / 64;
    // groups of kTrainWaves tiles; every wave runs every group (the block
    // synchronises inside the backward pass), a tile past the end as 64 invalid records
    for (uint32_t group = blockIdx.x; group * kTrainWaves < tiles; group += gridDim.x) {
        const uint32_t jr = (group * kTrainWaves + wave) * 64 + lane;
        const bool first = group == blockIdx.x;
        const bool valid = jr < a.batch;
        Enc enc;
        float3 wi = f3(0.f, 1.f, 0.f);
        float w = 0.f;
        if (valid) {
            const uint32_t r = (a.start + jr) % a.n;
            enc = encode(xyz(a.pos[r]), xyz(a.nrm[r]), xyz(a.wo[r]), net.box);
            const float4 d = a.wi[r];
            wi = xyz(d);
            w = d.w * a.wscale;
        } else {
#pragma unroll
            for (int k = 0; k < 7; k++) enc.v[k] = 0.f;
        }
        // forward; the inputs of layers 0, 2 and 4 stay in registers (checkpoints),
        // those of layers 1 and 3 are recomputed during the backward pass
        u32x4 x0[2][2], x2[4][2], x4[4][2];
        input_operands(enc, x0);
        {
            u32x4 x1[4][2];
            layer_to<kIn>(net, 0, x0, x1);
            layer_to<kHid>(net, 1, x1, x2);
        }
        {
            u32x4 x3[4][2];
            layer_to<kHid>(net, 2, x2, x3);
            layer_to<kHid>(net, 3, x3, x4);
        }
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            f32x16 acc[2];
            tile_fwd<kHid>(net.fwd + fwd_offset(4), net.bias + 64 * 4, kOut, x4, mt, acc);
            float *sb = L.stage + p * kStageStride + 4 * h;
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int o0 = 32 * mt + drow(v, 0);
                if (o0 + 4 * h < kOut) {
                    sb[o0] = acc[0][v];
                    sb[32 * kStageStride + o0] = acc[1][v];
                }
            }
        }
        wave_sync_lds();
        // loss gradient of this lane's record, written over its outputs
        {
            float *row = L.stage + lane * kStageStride;
            const float l = loss_grad(row, wi, w, row);
            if (valid) loss += l;
            else
                for (int k = 0; k < kOut; k++) row[k] = 0.f;
        }
        wave_sync_lds();
        // dY_4 as packed operands: element e of dy[ks][nt] is output kperm(ks, h, e) of record 32 nt + p
        u32x4 dacc[4][2];
        {
            const float *sb = L.stage + p * kStageStride + 4 * h;
#pragma unroll
            for (int ks = 0; ks < 4; ks++)
#pragma unroll
                for (int nt = 0; nt < 2; nt++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        // output kperm(ks, h, e) = kperm(ks, 0, e) + 4 h; outputs >= 40 are padding
                        const int o0 = kperm(ks, 0, 2 * i), o1 = kperm(ks, 0, 2 * i + 1);
                        const float v0 = o0 + 4 * h < kOut ? sb[32 * nt * kStageStride + o0] : 0.f;
                        const float v1 = o1 + 4 * h < kOut ? sb[32 * nt * kStageStride + o1] : 0.f;
                        dacc[ks][nt][i] = pack2(v0, v1);
                    }
        }
        wave_sync_lds();  // the stage is read before the transposed tiles overwrite it
        backward_layer<4>(net, wl, grad, first, bpart, x4, dacc);
        {
            u32x4 x3[4][2];
            layer_to<kHid>(net, 2, x2, x3);
            backward_layer<3>(net, wl, grad, first, bpart, x3, dacc);
        }
        backward_layer<2>(net, wl, grad, first, bpart, x2, dacc);
        {
            u32x4 x1[4][2];
            layer_to<kIn>(net, 0, x0, x1);
            backward_layer<1>(net, wl, grad, first, bpart, x1, dacc);
        }
        backward_layer<0>(net, wl, grad, first, bpart, x0, dacc);
    }
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
    if (lane == 0) wave_loss[wave] = loss;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.f;
        for (int k = 0; k < kTrainWaves; k++) s += wave_loss[k];
        a.loss[blockIdx.x] = s;
    }
}

// ---------------------------------------------------------------- adam
struct AdamArgs {
    float *params, *m, *v;
    const float *partials;
    int blocks;
    float lr_t;  // lr * sqrt(1 - beta2^t) / (1 - beta1^t)
    float beta1, beta2, eps_t;  // eps * sqrt(1 - beta2^t)
    float *grad_out;            // non-null: store the summed gradient, no update
    float *avg;                 // non-null: avg = decay avg + (1 - decay) params after the update
    float avg_decay;
};
// 64 parameters per block, 16 threads per parameter: each sums every 16th
// block partial (coalesced over the parameters, four loads in flight), then
// one thread per parameter adds the 16 sums in a fixed order.
constexpr int kAdamParams = 64, kAdamRows = 16, kAdamBlock = kAdamParams * kAdamRows;
__global__ __launch_bounds__(kAdamBlock) void k_adam(AdamArgs a) {
    __shared__ float part[kAdamBlock];
    const int pl = threadIdx.x % kAdamParams, q = threadIdx.x / kAdamParams;
    const int j = blockIdx.x * kAdamParams + pl;
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    if (j < kNumParams) {
        const float *col = a.partials + j;
        int b = q;
        for (; b + 3 * kAdamRows < a.blocks; b += 4 * kAdamRows)
#pragma unroll
            for (int u = 0; u < 4; u++) s[u] += col[size_t(b + u * kAdamRows) * kNumParams];
        for (; b < a.blocks; b += kAdamRows) s[0] += col[size_t(b) * kNumParams];
    }
    part[threadIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
    __syncthreads();
    if (q != 0 || j >= kNumParams) return;
    float g = 0.f;
#pragma unroll
    for (int r = 0; r < kAdamRows; r++) g += part[r * kAdamParams + pl];
    if (a.grad_out) {
        a.grad_out[j] = g;
        return;
    }
    const float m = a.beta1 * a.m[j] + (1.f - a.beta1) * g;
    const float v = a.beta2 * a.v[j] + (1.f - a.beta2) * g * g;
    a.m[j] = m;
    a.v[j] = v;
    const float p = a.params[j] - a.lr_t * m / (sqrtf(v) + a.eps_t);
    a.params[j] = p;
    if (a.avg) a.avg[j] = a.avg_decay * a.avg[j] + (1.f - a.avg_decay) * p;
}

// ---------------------------------------------------------------- finalize
// weight = mean over channels of (final radiance - radiance before the
// successor vertex) / throughput before the vertex = f cos L_i / p(wi).
__global__ void k_finalize(Records rec, uint32_t n, const float4 *sample_rad, double *wsum) {
    float s = 0.f;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        const uint32_t sidx = __float_as_uint(rec.pos[r].w);
        const float4 fin = sample_rad[sidx];
        const float4 base = rec.rad[r];
        const float4 t = rec.thr[r];
        float w = 0.f;
        w += t.x > 0.f ? fmaxf(fin.x - base.x, 0.f) / t.x : 0.f;
        w += t.y > 0.f ? fmaxf(fin.y - base.y, 0.f) / t.y : 0.f;
        w += t.z > 0.f ? fmaxf(fin.z - base.z, 0.f) / t.z : 0.f;
        w *= 1.f / 3.f;
        if (!(w < 3.0e38f)) w = 0.f;
        rec.wi[r].w = w;
        s += w;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (__lane_id() == 0 && s != 0.f) atomicAdd(wsum, double(s));
}

// ---------------------------------------------------------------- tests
template <bool FP8>
__global__ __launch_bounds__(kTestBlock) void k_forward_test(Net net, const float *pos, const float *nrm, const float *wo,
                                                             uint32_t n, float *out) {
    const uint32_t r = blockIdx.x * 64 + threadIdx.x;
    Enc e;
    if (r < n)
        e = encode(f3(pos[3 * r], pos[3 * r + 1], pos[3 * r + 2]), f3(nrm[3 * r], nrm[3 * r + 1], nrm[3 * r + 2]),
                   f3(wo[3 * r], wo[3 * r + 1], wo[3 * r + 2]), net.box);
    else
        for (int k = 0; k < 7; k++) e.v[k] = 0.f;
    float o[kOut];
    if (FP8) forward8(net, e, o);
    else forward(net, e, o);
    if (r < n)
        for (int k = 0; k < kOut; k++) out[size_t(r) * kOut + k] = o[k];
}

__global__ __launch_bounds__(kTestBlock) void k_mixture_test(const float *raw, const float *dirs, const float *u, uint32_t n,
                                                             float *pdf_dirs, float *samples, float *sample_pdf) {
    __shared__ float stage[64 * kStageStride];
    const uint32_t r = blockIdx.x * 64 + threadIdx.x;
    if (r >= n) return;
    float *L = stage + threadIdx.x * kStageStride;
    for (int k = 0; k < kOut; k++) L[k] = raw[size_t(r) * kOut + k];
    float w[kLobes];
    decode(L, w);
    pdf_dirs[r] = pdf(L, f3(dirs[3 * r], dirs[3 * r + 1], dirs[3 * r + 2]));
    const float3 s = sample(L, w, u[3 * r], u[3 * r + 1], u[3 * r + 2]);
    samples[3 * r] = s.x;
    samples[3 * r + 1] = s.y;
    samples[3 * r + 2] = s.z;
    sample_pdf[r] = pdf(L, s);
}

}  // namespace

// ============================================================== host
struct wr_guide {
    wr_scene *scene = nullptr;
    wr_guide_config cfg{};
    float *d_params = nullptr, *d_m = nullptr, *d_v = nullptr;
    // exponential average of the parameters after each Adam step (cfg.avg_decay),
    // zero-started at creation and set_params and debiased by 1 / (1 - decay^avg_steps):
    // the renderer's inference copies are packed from it
    float *d_avg = nullptr;
    uint64_t avg_steps = 0;
    u32x4 *d_fwd = nullptr, *d_bwd = nullptr;
    float *d_bias = nullptr;  // packed biases (kBiasPacked)
    u32x2 *d_fwd8 = nullptr;  // e4m3 forward operands (inference variant)
    float *d_wscale = nullptr;  // their per-layer scales
    float *d_partials = nullptr, *d_loss = nullptr;
    double *d_wsum = nullptr;
    Records recs[2]{};             // renders: pass q writes set q % 2 while set (q - 1) % 2 trains
    void *rec_pool = nullptr;
    // Training of a render pass runs on tstream; the host goes on enqueueing the
    // next pass, whose inference waits for it on the GPU.  Inference reads its
    // own copies of the packed operands: set k holds the network after the
    // training of the last pass that wrote records set k.
    hipStream_t tstream = nullptr;
    u32x4 *d_ifwd[2] = {nullptr, nullptr};
    float *d_ibias[2] = {nullptr, nullptr};
    u32x2 *d_ifwd8[2] = {nullptr, nullptr};
    float *d_iwscale[2] = {nullptr, nullptr};
    void *infer_pool = nullptr;
    hipEvent_t ev_trained[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};
    std::vector<std::pair<hipEvent_t, hipEvent_t>> train_marks;  // per trained pass: start, end (tstream)
    int last_blocks = 0;          // blocks of the last render training step (its loss partials)
    float4 *d_draw_dir = nullptr;  // per path-pool entry: the inference kernel's direction and q there
    float *d_draw_qnee = nullptr;  // and q toward the light sample (same allocation)
    uint32_t draw_capacity = 0;
    uint64_t step = 0;
    int train_blocks = 0;
    double last_loss = 0.0;
    // adaptive record selection: recorded vertices per selected path
    double verts_per_path = 4.0;
};

namespace {

// splitmix64, for the weight initialisation
struct HostRng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return double(next() >> 11) * 0x1p-53; }
};

// He-uniform hidden layers; a small output layer whose biases start the
// mixture as 8 equal lobes of sharpness 2 along the cube diagonals.
void init_params(uint32_t seed, std::vector<float> &p) {
    p.assign(kNumParams, 0.f);
    HostRng rng{0x243f6a8885a308d3ull ^ seed};
    for (int l = 0; l < kLayers; l++) {
        const int K = layer_in(l), N = layer_out(l);
        const double lim = std::sqrt(6.0 / K) * (l == kLayers - 1 ? 0.1 : 1.0);
        for (int j = 0; j < K * N; j++) p[w_offset(l) + j] = float((2.0 * rng.uniform() - 1.0) * lim);
    }
    float *b = p.data() + b_offset(kLayers - 1);
    for (int k = 0; k < kLobes; k++) {
        b[5 * k + 0] = 0.f;
        b[5 * k + 1] = (k & 1) ? 1.f : -1.f;
        b[5 * k + 2] = (k & 2) ? 1.f : -1.f;
        b[5 * k + 3] = (k & 4) ? 1.f : -1.f;
        b[5 * k + 4] = float(std::log(std::expm1(2.0)));  // softplus^-1(2)
    }
}

Net net_view(const wr_guide *g) {
    Net n;
    n.fwd = g->d_fwd;
    n.bwd = g->d_bwd;
    n.bias = g->d_bias;
    n.fwd8 = g->d_fwd8;
    n.wscale = g->d_wscale;
    for (int k = 0; k < 3; k++) {
        n.box[k] = g->scene->box[k];
        n.box[3 + k] = float(1.0 / std::max(double(g->scene->box[3 + k]) - double(g->scene->box[k]), 1e-20));
    }
    return n;
}

int pack(wr_guide *g, hipStream_t st) {
    const int n = kFwdOps + kBwdOps;
    k_pack<<<kLayers + (n + kPack8Block - 1) / kPack8Block, kPack8Block, 0, st>>>(g->d_params, 1.f, g->d_fwd, g->d_bwd,
                                                                                    g->d_bias, g->d_fwd8, g->d_wscale);
    WR_HIP(hipGetLastError());
    return WR_OK;
}

// Inference copy `set` <- the debiased parameter average (the parameters
// themselves before the first averaged step, or with averaging off).
int pack_inference(wr_guide *g, int set, hipStream_t st) {
    const bool avg = g->cfg.avg_decay > 0.f && g->avg_steps > 0;
    const float sc = avg ? float(1.0 / (1.0 - std::pow(double(g->cfg.avg_decay), double(g->avg_steps)))) : 1.f;
    k_pack<<<kLayers + (kFwdOps + kPack8Block - 1) / kPack8Block, kPack8Block, 0, st>>>(
        avg ? g->d_avg : g->d_params, sc, g->d_ifwd[set], nullptr, g->d_ibias[set], g->d_ifwd8[set], g->d_iwscale[set]);
    WR_HIP(hipGetLastError());
    return WR_OK;
}

int reset_adam(wr_guide *g) {
    WR_HIP(hipMemsetAsync(g->d_m, 0, sizeof(float) * kNumParams, g->scene->stream));
    WR_HIP(hipMemsetAsync(g->d_v, 0, sizeof(float) * kNumParams, g->scene->stream));
    WR_HIP(hipMemsetAsync(g->d_avg, 0, sizeof(float) * kNumParams, g->scene->stream));
    g->step = 0;
    g->avg_steps = 0;
    return WR_OK;
}

// One training step (or, with grad_out, the gradient only) over `batch`
// records starting at `start` of the n records in recs.
int train_step(wr_guide *g, const float4 *pos, const float4 *nrm, const float4 *wo, const float4 *wi, uint32_t n,
               uint32_t start, uint32_t batch, float wscale, float *grad_out, double *loss, hipStream_t st) {
    TrainArgs a;
    a.pos = pos; a.nrm = nrm; a.wo = wo; a.wi = wi;
    a.n = n; a.start = start; a.batch = batch; a.wscale = wscale;
    a.partials = g->d_partials;
    a.loss = g->d_loss;
    const int tiles = int((batch + 63) / 64);
    const int blocks = std::max(1, std::min((tiles + kTrainWaves - 1) / kTrainWaves, g->train_blocks));
    k_train<<<blocks, 64 * kTrainWaves, 0, st>>>(net_view(g), a);
    WR_HIP(hipGetLastError());
    g->last_blocks = blocks;
    AdamArgs ad;
    ad.params = g->d_params; ad.m = g->d_m; ad.v = g->d_v;
    ad.partials = g->d_partials;
    ad.blocks = blocks;
    ad.grad_out = grad_out;
    ad.avg = grad_out || !(g->cfg.avg_decay > 0.f) ? nullptr : g->d_avg;
    ad.avg_decay = g->cfg.avg_decay;
    if (ad.avg) g->avg_steps++;
    if (!grad_out) {
        g->step++;
        const double t = double(g->step);
        const double c1 = 1.0 - std::pow(double(g->cfg.beta1), t), c2 = 1.0 - std::pow(double(g->cfg.beta2), t);
        ad.lr_t = float(g->cfg.learning_rate * std::sqrt(c2) / c1);
        ad.eps_t = float(g->cfg.epsilon * std::sqrt(c2));
    } else {
        ad.lr_t = 0.f;
        ad.eps_t = 1.f;
    }
    ad.beta1 = g->cfg.beta1;
    ad.beta2 = g->cfg.beta2;
    k_adam<<<(kNumParams + kAdamParams - 1) / kAdamParams, kAdamBlock, 0, st>>>(ad);
    WR_HIP(hipGetLastError());
    if (!grad_out) {
        int rc = pack(g, st);
        if (rc) return rc;
    }
    if (loss) {
        std::vector<float> l(blocks);
        WR_HIP(hipMemcpyAsync(l.data(), g->d_loss, sizeof(float) * blocks, hipMemcpyDeviceToHost, st));
        WR_HIP(hipStreamSynchronize(st));
        double s = 0.0;
        for (float x : l) s += x;
        *loss = s;
    }
    return WR_OK;
}

// Host records -> temporary device arrays (float4 each).
struct DeviceRecords {
    float4 *buf = nullptr;
    float4 *pos = nullptr, *nrm = nullptr, *wo = nullptr, *wi = nullptr;
    ~DeviceRecords() { if (buf) (void)hipFree(buf); }
    int upload(hipStream_t st, const float *p, const float *nr, const float *o, const float *i, const float *w, uint32_t n) {
        std::vector<float4> h(4 * size_t(n));
        for (uint32_t r = 0; r < n; r++) {
            h[r] = make_float4(p[3 * r], p[3 * r + 1], p[3 * r + 2], 0.f);
            h[n + r] = make_float4(nr[3 * r], nr[3 * r + 1], nr[3 * r + 2], 0.f);
            h[2 * size_t(n) + r] = make_float4(o[3 * r], o[3 * r + 1], o[3 * r + 2], 0.f);
            h[3 * size_t(n) + r] = make_float4(i[3 * r], i[3 * r + 1], i[3 * r + 2], w[r]);
        }
        WR_HIP(hipMalloc(reinterpret_cast<void **>(&buf), sizeof(float4) * h.size()));
        WR_HIP(hipMemcpyAsync(buf, h.data(), sizeof(float4) * h.size(), hipMemcpyHostToDevice, st));
        WR_HIP(hipStreamSynchronize(st));
        pos = buf; nrm = buf + n; wo = buf + 2 * size_t(n); wi = buf + 3 * size_t(n);
        return WR_OK;
    }
};

}  // namespace

// ---------------------------------------------------------------- internal interface for the renderer
// A render call: guide_render_begin, then per pass q (record set k = q % 2)
// guide_pass_net (the network trained on pass q - 1), guide_begin_pass,
// guide_shade_view, the pass, guide_end_pass (finalise on the render stream,
// train on tstream; the host does not wait), and guide_render_end (waits for
// the last training).
int guide_render_begin(wr_guide *g) {
    hipStream_t st = g->scene->stream;
    WR_HIP(hipStreamSynchronize(g->tstream));
    for (int k = 0; k < 2; k++) {
        int rc = pack_inference(g, k, st);
        if (rc) return rc;
        g->pending[k] = false;
    }
    g->train_marks.clear();
    return WR_OK;
}

// Pass q (records set = q % 2) renders with the network trained on pass q - 1
// (inference copy set ^ 1): the render stream waits for that training on the
// GPU, which also orders it after the training of pass q - 2 that read this
// pass's record set.  With `overlap`, pass q renders with the network trained
// on pass q - 2 (inference copy `set`) and waits only for that training,
// which read this pass's record set: the training on pass q - 1 then runs on
// its own stream beside the pass.  (With passes of 8 spp and 32-64 steps the
// stale network cost more variance than the overlap saved; with 2-spp passes
// and 256 steps the training is as long as the pass.)
int guide_pass_net(wr_guide *g, int set, bool overlap, guide::Net *net) {
    const int use = overlap ? set : set ^ 1;
    if (g->pending[use]) WR_HIP(hipStreamWaitEvent(g->scene->stream, g->ev_trained[use], 0));
    g->pending[use] = false;
    *net = net_view(g);
    net->fwd = g->d_ifwd[use];
    net->bias = g->d_ibias[use];
    net->fwd8 = g->d_ifwd8[use];
    net->wscale = g->d_iwscale[use];
    net->bwd = nullptr;
    return WR_OK;
}

int guide_shade_view(wr_guide *g, int set, float alpha, float record_prob, bool collect, uint32_t max_bounce,
                     guide::Shade *out) {
    out->draw.dir = g->d_draw_dir;
    out->draw.q_nee = g->d_draw_qnee;
    out->rec = g->recs[set];
    out->alpha = alpha;
    out->record_prob = record_prob;
    out->collect = collect ? 1 : 0;
    out->max_bounce = max_bounce;
    return WR_OK;
}

int guide_begin_pass(wr_guide *g, int set, uint64_t samples, float *record_prob) {
    k_fill_u32_guide<<<1, 64, 0, g->scene->stream>>>(g->recs[set].count, 1);
    WR_HIP(hipGetLastError());
    const double want = 0.9 * double(g->recs[set].capacity);
    const double expect = double(samples) * g->verts_per_path;
    *record_prob = float(std::min(1.0, want / std::max(expect, 1.0)));
    return WR_OK;
}

// After a pass: weights from the final radiance (render stream), then `steps`
// Adam steps on tstream, which also leave the trained network in inference set
// `set`.
int guide_end_pass(wr_guide *g, int set, const float4 *sample_rad, uint64_t samples, float record_prob, uint32_t steps,
                   wr_guide_stats *gs) {
    hipStream_t st = g->scene->stream, ts = g->tstream;
    const Records &rec = g->recs[set];
    uint32_t count = 0;
    WR_HIP(hipMemcpyAsync(&count, rec.count, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    WR_HIP(hipStreamSynchronize(st));
    const double selected = double(samples) * double(record_prob);
    if (selected >= 1.0 && count > 0) g->verts_per_path = std::max(1.0, double(count) / selected);
    const uint32_t n = std::min(count, rec.capacity);
    gs->records += n;
    if (n == 0 || steps == 0) return WR_OK;
    k_fill_f64_guide<<<1, 64, 0, st>>>(g->d_wsum, 1);
    k_finalize<<<std::min<uint32_t>((n + 255) / 256, 4096), 256, 0, st>>>(rec, n, sample_rad, g->d_wsum);
    WR_HIP(hipGetLastError());
    double wsum = 0.0;
    WR_HIP(hipMemcpyAsync(&wsum, g->d_wsum, sizeof(double), hipMemcpyDeviceToHost, st));
    WR_HIP(hipStreamSynchronize(st));  // the records are final
    if (!(wsum > 0.0)) return WR_OK;
    hipEvent_t t0, t1;
    WR_HIP(hipEventCreate(&t0));
    WR_HIP(hipEventCreate(&t1));
    g->train_marks.emplace_back(t0, t1);
    WR_HIP(hipEventRecord(t0, ts));
    const uint32_t batch = std::min(n, g->cfg.batch_size);
    // mean weight 1 over the pass's records; loss averaged over the batch
    const float wscale = float(double(n) / wsum / double(batch));
    for (uint32_t k = 0; k < steps; k++) {
        const uint32_t start = uint32_t((uint64_t(k) * batch) % n);
        int rc = train_step(g, rec.pos, rec.nrm, rec.wo, rec.wi, n, start, batch, wscale, nullptr, nullptr, ts);
        if (rc) return rc;
        gs->train_steps++;
    }
    int rc = pack_inference(g, set, ts);
    if (rc) return rc;
    WR_HIP(hipEventRecord(t1, ts));
    WR_HIP(hipEventRecord(g->ev_trained[set], ts));
    g->pending[set] = true;
    return WR_OK;
}

// Waits for the last training, then the loss of the last step and the
// training time.
int guide_render_end(wr_guide *g, wr_guide_stats *gs) {
    WR_HIP(hipStreamSynchronize(g->tstream));
    for (auto &m : g->train_marks) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, m.first, m.second) == hipSuccess) gs->train_ms += ms;
        (void)hipEventDestroy(m.first);
        (void)hipEventDestroy(m.second);
    }
    if (!g->train_marks.empty() && g->last_blocks > 0) {
        std::vector<float> l(g->last_blocks);
        WR_HIP(hipMemcpy(l.data(), g->d_loss, sizeof(float) * l.size(), hipMemcpyDeviceToHost));
        double sum = 0.0;
        for (float x : l) sum += x;
        gs->last_loss = sum;
    }
    g->train_marks.clear();
    g->pending[0] = g->pending[1] = false;
    return WR_OK;
}


int guide_reserve_draws(wr_guide *g, uint32_t capacity) {
    if (g->draw_capacity >= capacity) return WR_OK;
    WR_HIP(hipStreamSynchronize(g->scene->stream));
    if (g->d_draw_dir) WR_HIP(hipFree(g->d_draw_dir));
    g->d_draw_dir = nullptr;
    g->d_draw_qnee = nullptr;
    g->draw_capacity = 0;
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&g->d_draw_dir), (sizeof(float4) + sizeof(float)) * size_t(capacity)));
    g->d_draw_qnee = reinterpret_cast<float *>(g->d_draw_dir + capacity);
    g->draw_capacity = capacity;
    return WR_OK;
}
wr_scene *guide_scene(const wr_guide *g) { return g->scene; }

// ---------------------------------------------------------------- C-ABI
void wr_guide_config_default(wr_guide_config *c) {
    if (!c) return;
    c->learning_rate = 1e-2f;
    c->beta1 = 0.9f;
    c->beta2 = 0.999f;
    c->epsilon = 1e-8f;
    c->batch_size = 1u << 18;
    c->record_capacity = 1u << 20;
    c->seed = 0;
    c->avg_decay = 0.99f;
}

void wr_guide_destroy(wr_guide *g) {
    if (!g) return;
    (void)hipSetDevice(g->scene->dev->ordinal);
    (void)hipStreamSynchronize(g->scene->stream);
    if (g->tstream) (void)hipStreamSynchronize(g->tstream);
    for (auto &m : g->train_marks) {
        (void)hipEventDestroy(m.first);
        (void)hipEventDestroy(m.second);
    }
    for (int k = 0; k < 2; k++) {
        if (g->ev_trained[k]) (void)hipEventDestroy(g->ev_trained[k]);
    }
    if (g->tstream) (void)hipStreamDestroy(g->tstream);
    for (void *p : {(void *)g->d_params, (void *)g->d_m, (void *)g->d_v, (void *)g->d_avg, (void *)g->d_fwd, (void *)g->d_bwd, (void *)g->d_bias,
                    (void *)g->d_fwd8, (void *)g->d_wscale, (void *)g->d_partials, (void *)g->d_loss, (void *)g->d_wsum,
                    g->rec_pool, g->infer_pool, (void *)g->d_draw_dir})
        if (p) (void)hipFree(p);
    delete g;
}

static int guide_create_impl(wr_guide *g) {
    int rc = set_device(g->scene->dev);
    if (rc) return rc;
    int ndev_blocks = g->scene->dev->n_cu;
    g->train_blocks = std::min(kMaxTrainBlocks, ndev_blocks);
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&g->d_params), sizeof(float) * kNumParams));
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&g->d_m), sizeof(float) * kNumParams));
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&g->d_v), sizeof(float) * kNumParams));
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&g->d_avg), sizeof(float) * kNumParams));
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&g->d_fwd), sizeof(u32x4) * kFwdOps));
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&g->d_bwd), sizeof(u32x4) * kBwdOps));
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&g->d_bias), sizeof(float) * kBiasPacked));
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&g->d_fwd8), sizeof(u32x2) * kFwdOps));
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&g->d_wscale), sizeof(float) * kLayers));
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&g->d_partials), sizeof(float) * kNumParams * size_t(g->train_blocks)));
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&g->d_loss), sizeof(float) * g->train_blocks));
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&g->d_wsum), sizeof(double)));
    const size_t cap = g->cfg.record_capacity;
    WR_HIP(hipMalloc(&g->rec_pool, 2 * (sizeof(float4) * 6 * cap + 256)));
    for (int k = 0; k < 2; k++) {
        float4 *b = reinterpret_cast<float4 *>(static_cast<char *>(g->rec_pool) + k * (sizeof(float4) * 6 * cap + 256));
        Records &r = g->recs[k];
        r.pos = b;
        r.nrm = b + cap;
        r.wo = b + 2 * cap;
        r.wi = b + 3 * cap;
        r.thr = b + 4 * cap;
        r.rad = b + 5 * cap;
        r.count = reinterpret_cast<uint32_t *>(b + 6 * cap);
        r.capacity = uint32_t(cap);
    }
    // inference copies of the packed operands, two sets, 256-byte aligned pieces
    const size_t fwd_b = (sizeof(u32x4) * kFwdOps + 255) / 256 * 256, bias_b = (sizeof(float) * kBiasPacked + 255) / 256 * 256;
    const size_t fwd8_b = (sizeof(u32x2) * kFwdOps + 255) / 256 * 256, ws_b = 256;
    const size_t set_b = fwd_b + bias_b + fwd8_b + ws_b;
    WR_HIP(hipMalloc(&g->infer_pool, 2 * set_b));
    for (int k = 0; k < 2; k++) {
        char *b = static_cast<char *>(g->infer_pool) + k * set_b;
        g->d_ifwd[k] = reinterpret_cast<u32x4 *>(b);
        g->d_ibias[k] = reinterpret_cast<float *>(b + fwd_b);
        g->d_ifwd8[k] = reinterpret_cast<u32x2 *>(b + fwd_b + bias_b);
        g->d_iwscale[k] = reinterpret_cast<float *>(b + fwd_b + bias_b + fwd8_b);
        WR_HIP(hipEventCreateWithFlags(&g->ev_trained[k], hipEventDisableTiming));
    }
    WR_HIP(hipStreamCreateWithFlags(&g->tstream, hipStreamNonBlocking));
    std::vector<float> p;
    init_params(g->cfg.seed, p);
    return wr_guide_set_params(g, p.data());
}

int wr_guide_create(wr_scene *s, const wr_guide_config *cfg, wr_guide **out) {
    if (!s || !out) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    wr_guide_config c;
    wr_guide_config_default(&c);
    if (cfg) c = *cfg;
    if (c.batch_size == 0 || c.record_capacity == 0 || c.record_capacity > (1u << 26))
        return set_error(WR_ERR_INVALID_ARGUMENT, "batch_size must be > 0 and record_capacity in [1, 2^26]");
    if (!(c.learning_rate > 0.f) || !(c.beta1 >= 0.f && c.beta1 < 1.f) || !(c.beta2 >= 0.f && c.beta2 < 1.f) ||
        !(c.epsilon > 0.f) || !(c.avg_decay >= 0.f && c.avg_decay < 1.f))
        return set_error(WR_ERR_INVALID_ARGUMENT, "bad Adam hyper-parameters");
    wr_guide *g = new wr_guide;
    g->scene = s;
    g->cfg = c;
    int rc = guide_create_impl(g);
    if (rc) {
        std::string msg = wr::error_string();
        wr_guide_destroy(g);
        wr::error_string() = msg;
        return rc;
    }
    *out = g;
    return WR_OK;
}

int wr_guide_get_params(wr_guide *g, float *params) {
    if (!g || !params) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    int rc = set_device(g->scene->dev);
    if (rc) return rc;
    WR_HIP(hipMemcpyAsync(params, g->d_params, sizeof(float) * kNumParams, hipMemcpyDeviceToHost, g->scene->stream));
    WR_HIP(hipStreamSynchronize(g->scene->stream));
    return WR_OK;
}

int wr_guide_set_params(wr_guide *g, const float *params) {
    if (!g || !params) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    int rc = set_device(g->scene->dev);
    if (rc) return rc;
    for (int j = 0; j < kNumParams; j++)
        if (!std::isfinite(params[j])) return set_error(WR_ERR_INVALID_ARGUMENT, "non-finite parameter");
    WR_HIP(hipMemcpyAsync(g->d_params, params, sizeof(float) * kNumParams, hipMemcpyHostToDevice, g->scene->stream));
    rc = reset_adam(g);
    if (!rc) rc = pack(g, g->scene->stream);
    if (rc) return rc;
    WR_HIP(hipStreamSynchronize(g->scene->stream));
    return WR_OK;
}

static int check_records(const float *pos, const float *nrm, const float *wo, uint32_t n) {
    if (n > 0 && (!pos || !nrm || !wo)) return set_error(WR_ERR_INVALID_ARGUMENT, "null record arrays");
    if (n > (1u << 26)) return set_error(WR_ERR_UNSUPPORTED, "too many records (limit 2^26)");
    return WR_OK;
}

static int forward_impl(wr_guide *g, const float *pos, const float *nrm, const float *wo, uint32_t n, float *out,
                        bool fp8) {
    if (!g || (n && !out)) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    int rc = check_records(pos, nrm, wo, n);
    if (!rc) rc = set_device(g->scene->dev);
    if (rc || n == 0) return rc;
    hipStream_t st = g->scene->stream;
    float *d = nullptr;
    const size_t in_bytes = sizeof(float) * 9 * size_t(n), out_bytes = sizeof(float) * kOut * size_t(n);
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&d), in_bytes + out_bytes));
    float *dp = d, *dn = d + 3 * size_t(n), *dw = d + 6 * size_t(n), *dout = d + 9 * size_t(n);
    hipError_t e = hipMemcpyAsync(dp, pos, in_bytes / 3, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dn, nrm, in_bytes / 3, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dw, wo, in_bytes / 3, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        if (fp8) k_forward_test<true><<<(n + 63) / 64, kTestBlock, 0, st>>>(net_view(g), dp, dn, dw, n, dout);
        else k_forward_test<false><<<(n + 63) / 64, kTestBlock, 0, st>>>(net_view(g), dp, dn, dw, n, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, out_bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d);
    WR_HIP(e);
    return WR_OK;
}

int wr_guide_forward(wr_guide *g, const float *pos, const float *nrm, const float *wo, uint32_t n, float *out) {
    return forward_impl(g, pos, nrm, wo, n, out, false);
}

int wr_guide_forward_fp8(wr_guide *g, const float *pos, const float *nrm, const float *wo, uint32_t n, float *out) {
    return forward_impl(g, pos, nrm, wo, n, out, true);
}

int wr_guide_gradient(wr_guide *g, const float *pos, const float *nrm, const float *wo, const float *wi,
                      const float *weight, uint32_t n, float *grad, double *loss) {
    if (!g || !grad || !wi || !weight) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    int rc = check_records(pos, nrm, wo, n);
    if (!rc) rc = set_device(g->scene->dev);
    if (rc) return rc;
    if (n == 0) return set_error(WR_ERR_INVALID_ARGUMENT, "no records");
    DeviceRecords dr;
    rc = dr.upload(g->scene->stream, pos, nrm, wo, wi, weight, n);
    if (rc) return rc;
    float *dg = nullptr;
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&dg), sizeof(float) * kNumParams));
    rc = train_step(g, dr.pos, dr.nrm, dr.wo, dr.wi, n, 0, n, 1.f / float(n), dg, loss, g->scene->stream);
    hipError_t e = hipSuccess;
    if (!rc) {
        e = hipMemcpyAsync(grad, dg, sizeof(float) * kNumParams, hipMemcpyDeviceToHost, g->scene->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g->scene->stream);
    }
    (void)hipFree(dg);
    if (rc) return rc;
    WR_HIP(e);
    return WR_OK;
}

int wr_guide_train(wr_guide *g, const float *pos, const float *nrm, const float *wo, const float *wi,
                   const float *weight, uint32_t n, uint32_t steps, double *loss) {
    if (!g || !wi || !weight) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    int rc = check_records(pos, nrm, wo, n);
    if (!rc) rc = set_device(g->scene->dev);
    if (rc) return rc;
    if (n == 0) return set_error(WR_ERR_INVALID_ARGUMENT, "no records");
    DeviceRecords dr;
    rc = dr.upload(g->scene->stream, pos, nrm, wo, wi, weight, n);
    if (rc) return rc;
    for (uint32_t k = 0; k < steps; k++) {
        rc = train_step(g, dr.pos, dr.nrm, dr.wo, dr.wi, n, 0, n, 1.f / float(n), nullptr,
                        (k + 1 == steps) ? loss : nullptr, g->scene->stream);
        if (rc) return rc;
    }
    WR_HIP(hipStreamSynchronize(g->scene->stream));
    return WR_OK;
}

int wr_guide_mixture(wr_guide *g, const float *raw, const float *dirs, const float *u, uint32_t n, float *pdf_dirs,
                     float *samples, float *sample_pdf) {
    if (!g || (n && (!raw || !dirs || !u || !pdf_dirs || !samples || !sample_pdf)))
        return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    if (n > (1u << 26)) return set_error(WR_ERR_UNSUPPORTED, "too many rows (limit 2^26)");
    int rc = set_device(g->scene->dev);
    if (rc || n == 0) return rc;
    hipStream_t st = g->scene->stream;
    float *d = nullptr;
    const size_t N = n;
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&d), sizeof(float) * N * (kOut + 3 + 3 + 1 + 3 + 1)));
    float *draw = d, *ddir = d + kOut * N, *du = ddir + 3 * N, *dpd = du + 3 * N, *ds = dpd + N, *dsp = ds + 3 * N;
    hipError_t e = hipMemcpyAsync(draw, raw, sizeof(float) * kOut * N, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(ddir, dirs, sizeof(float) * 3 * N, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(du, u, sizeof(float) * 3 * N, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        k_mixture_test<<<(n + 63) / 64, kTestBlock, 0, st>>>(draw, ddir, du, n, dpd, ds, dsp);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(pdf_dirs, dpd, sizeof(float) * N, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(samples, ds, sizeof(float) * 3 * N, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(sample_pdf, dsp, sizeof(float) * N, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d);
    WR_HIP(e);
    return WR_OK;
}
