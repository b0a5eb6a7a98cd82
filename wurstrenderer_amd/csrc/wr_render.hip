// wr_render.hip -- wavefront NEE + MIS path tracer for gfx950 and the C-ABI.
//
// Reference: PathNeeRenderer (src/renderer/pathnextevent.h:15-196) and the
// random walk of GeneralSubpathSampler (src/pathsampler/generalsubpathsampler.h:13-160).
// The reference's recursive per-sample loop over tiles becomes a wavefront
// over a fixed pool of paths, one iteration per path vertex:
//
//   regen     start the next unclaimed samples behind the continuing paths:
//             light vertex, camera ray, light->camera (t = 1) connection
//   extend    closest hit for every active path                (BVH4 traversal)
//   shade     emitter hit with MIS, NEE with MIS, BSDF sample; continuing paths
//             are compacted into the next state, a finished path stores its
//             radiance under its sample index
//   shadow    any hit for NEE and t = 1 rays; add visible contributions
//   rotate    the next state becomes current (device-side counters)
//   (per pass) accumulate: per-pixel FP64 sums of the samples' radiance
//
// Refilling the pool as paths end keeps every kernel's queue full until the
// last samples drain.  Path state is stored in queue order, so every kernel
// reads and writes it contiguously; compaction is one atomic per wave (ballot
// + mbcnt).  Every kernel reads its work count from device memory, and the
// host polls the counters only every few iterations.  Radiance is gathered in
// FP32 per path and summed per pixel in FP64 in sample order (reproducible);
// t = 1 splats go to the FP64 splat film with atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/wurst_amd.h"
#include "wr_device.h"
#include "wr_layout.h"
#include "wr_traverse.h"
#include "wr_guide.h"

using namespace wr;

#include "wr_host.h"

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t mbcnt64(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Wave-aggregated append: one atomic per wave.  All lanes must call it.
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool pred) {
    const unsigned long long mask = __ballot(pred);
    if (mask == 0) return kInvalid;
    const uint32_t leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return pred ? base + mbcnt64(mask) : kInvalid;
}

// Block-aggregated append for kShadeBlock-thread blocks: one atomic per block.
// All threads of the block must call it (it synchronises the block).
constexpr int kShadeBlock = 256;
constexpr int kWavesPerShadeBlock = kShadeBlock / 64;
__device__ __forceinline__ uint32_t block_append(uint32_t *counter, bool pred, uint32_t *lds /* kWavesPerShadeBlock+1 */) {
    const unsigned long long mask = __ballot(pred);
    const uint32_t wave = threadIdx.x / 64;
    if (lane_id() == 0) lds[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (int w = 0; w < kWavesPerShadeBlock; w++) {
            const uint32_t k = lds[w];
            lds[w] = total;
            total += k;
        }
        lds[kWavesPerShadeBlock] = total ? atomicAdd(counter, total) : 0u;
    }
    __syncthreads();
    const uint32_t pos = lds[kWavesPerShadeBlock] + lds[wave] + mbcnt64(mask);
    __syncthreads();  // lds is reused by the next call
    return pred ? pos : kInvalid;
}

__device__ __forceinline__ void wave_add_stat(unsigned long long *dst, uint32_t v) {
    uint32_t s = v;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane_id() == 0 && s) atomicAdd(dst, (unsigned long long)s);
}

__device__ __forceinline__ uint32_t active_count(const uint32_t *c) { return c[kQ] + c[kFresh]; }

// ============================================================== control kernels
// Number of new samples the next iteration starts: the free room of the pool
// or what is left of the pass.  Written here so regen needs no atomics.
__device__ __forceinline__ void plan_fresh(uint32_t *c, const RenderArgs &a) {
    const uint32_t head = c[kQ];
    const uint32_t left = a.total - min(c[kClaimed], a.total);
    const uint32_t fresh = min(a.capacity - head, left);
    c[kBase] = c[kClaimed];
    c[kFresh] = fresh;
    c[kClaimed] += fresh;
}
__global__ void k_init(PathBuffers pb, RenderArgs a) {
    if (threadIdx.x < kNumCounterLines) pb.counters[threadIdx.x * kCounterStride] = 0u;
    __syncthreads();
    if (threadIdx.x == 0) plan_fresh(pb.counters, a);
}
// End of an iteration: the next state becomes current, work counters restart,
// the shadow queue traced in this iteration is emptied (the one the shade
// kernel filled stays for the next traversal launch), and the next refill is
// planned.
__global__ void k_rotate(PathBuffers pb, RenderArgs a) {
    uint32_t *c = pb.counters;
    const uint32_t traced = pb.next.shadow_count;
    pb.stats[0] += active_count(c);
    pb.stats[1] += c[traced];
    c[kQ] = c[kQn];
    c[kQn] = 0;
    c[traced] = 0;
    for (int x = 0; x < kXcds; x++) {
        c[kFetchExtX + x * kCounterStride] = 0;
        c[kFetchShdX + x * kCounterStride] = 0;
    }
    plan_fresh(c, a);
}
// The traversal copy of the BVH4 (NodeQ), one node per thread.  Per axis the
// grid starts at the smallest lo of the node's children and its spacing is
// (largest hi - origin) / 255, raised by ulps until plane 255, decoded as the
// kernels decode it (fmaf(q, scale, origin)), reaches the largest hi.  A child's low plane is the
// largest q whose decoded value is <= its lo, its high plane the smallest q
// whose decoded value is >= its hi: q = 0 decodes to the origin (<= every lo)
// and q = 255 to at least the largest hi, so both exist and the decoded box
// contains the child's box.
__global__ void k_quantize_nodes(const Node4 *__restrict__ in, NodeQ *__restrict__ out, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Node4 &nd = in[k];
    const float *lo[3] = {nd.lo_x, nd.lo_y, nd.lo_z}, *hi[3] = {nd.hi_x, nd.hi_y, nd.hi_z};
    NodeQ q;
    float scale[3];
    for (int a = 0; a < 3; a++) {
        float org = FLT_MAX, top = -FLT_MAX;
        for (int c = 0; c < 4; c++)
            if (nd.child[c] != kEmptyChild) {
                org = fminf(org, lo[a][c]);
                top = fmaxf(top, hi[a][c]);
            }
        if (!(org <= top)) org = top = 0.f;  // no child
        float sc = fmaxf(float((double(top) - double(org)) / 255.0), FLT_MIN);
        while (fmaf(255.f, sc, org) < top) sc = nextafterf(sc, INFINITY);
        uint32_t wl = 0, wh = 0;
        for (int c = 0; c < 4; c++) {
            uint32_t ql = 255, qh = 0;  // empty slot
            if (nd.child[c] != kEmptyChild) {
                const float l = lo[a][c], h = hi[a][c];
                int a0 = int(fmin(255.0, fmax(0.0, floor((double(l) - org) / sc))));
                while (a0 > 0 && fmaf(float(a0), sc, org) > l) a0--;
                while (a0 < 255 && fmaf(float(a0 + 1), sc, org) <= l) a0++;
                int b0 = int(fmin(255.0, fmax(0.0, ceil((double(h) - org) / sc))));
                while (b0 < 255 && fmaf(float(b0), sc, org) < h) b0++;
                while (b0 > 0 && fmaf(float(b0 - 1), sc, org) >= h) b0--;
                ql = uint32_t(a0);
                qh = uint32_t(b0);
            }
            wl |= ql << (8 * c);
            wh |= qh << (8 * c);
        }
        q.org[a] = org;
        scale[a] = sc;
        q.q[2 * a] = wl;
        q.q[2 * a + 1] = wh;
    }
    q.scale_x = scale[0];
    q.scale_y = scale[1];
    q.scale_z = scale[2];
    for (int c = 0; c < 4; c++) q.child[c] = nd.child[c];
    out[k] = q;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
    for (size_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = v;
}
__global__ void k_fill_f64(double *p, size_t n) {
    for (size_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = 0.0;
}

// ============================================================== batches
// A batch renders several (scene, film, parameters) jobs through one path pool
// and one launch of every kernel per iteration: the pass's samples are the
// jobs' samples laid end to end (job-major), a path carries its job in the
// high bits of meta.z (bounce in the low 8), and every per-path lookup reads
// the job's scene.  Single-scene kernels (BATCH = false) take the scene as a
// kernel argument instead, and their job is always 0.
struct BatchJob {
    DScene s;
    uint32_t offset;        // first pass sample of the job
    uint32_t count;         // samples of the job in this pass = k * wh
    uint32_t wh, k;         // pixels, samples per pixel in this pass
    uint32_t first_sample;  // sample index of the pass's first spp
    uint32_t seed;
    int N, use_mis;
    double *film_base, *film_splat;
};
constexpr uint32_t kBounceBits = 8;
constexpr uint32_t kBounceMask = (1u << kBounceBits) - 1;
constexpr int kMaxBatchJobs = 1 << 16;
// The job a pass sample belongs to: the last job whose offset is <= g.
__device__ __forceinline__ uint32_t find_job(const BatchJob *jobs, uint32_t n_jobs, uint32_t g) {
    uint32_t lo = 0, hi = n_jobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].offset <= g) lo = mid; else hi = mid;
    }
    return lo;
}
// Geometry of one lane's scene for traversal.
struct Geo {
    const NodeQ *nodes;
    const float4 *tris;
};
// The same over the full-precision BVH4 (per-pixel kernels, see node_test).
struct Geo4 {
    const Node4 *nodes;
    const float4 *tris;
};

// Camera sample of path `key`: lens point and primary direction from its
// lens and pixel-jitter dimensions (lens_point_u / primary_dir_u).
__device__ __forceinline__ float3 lens_point(const DCamera &cam, uint32_t key) {
    return lens_point_u(cam, u01(key, kDimLens), u01(key, kDimLens + 1));
}
__device__ __forceinline__ float3 primary_dir(const DCamera &cam, uint32_t key, uint32_t px, uint32_t py, float3 lens,
                                              float &pdfW) {
    return primary_dir_u(cam, u01(key, kDimPixel), u01(key, kDimPixel + 1), px, py, lens, pdfW);
}

// ============================================================== regenerate
// Fills the pool behind the continuing paths with the next unclaimed samples:
// createLightSubpath + createCameraSubpath up to the first ray, with the t = 1
// light->camera connection (pathnextevent.h:53-99; thinlens.h:121-166), which
// PrimitivePathRenderer (a.nee == 0) does not make.  For LightSplattingRenderer
// (a.light) an entry is a light subpath instead (lightsplatting.h:58-108,
// generalsubpathsampler.h:185-268): its light vertex connects to the camera
// with weight 1, and its first ray leaves the light in a cosine-distributed
// direction (DiffuseAreaLight::sampleLe1CosByPdf, contribution 1).
template <bool BATCH, bool ENV = false>
__global__ __launch_bounds__(kShadeBlock) void k_regen(DScene s1, const BatchJob *__restrict__ jobs, uint32_t n_jobs,
                                                       PathBuffers pb, RenderArgs a) {
    __shared__ uint32_t agg[kWavesPerShadeBlock + 1];
    const uint32_t head = pb.counters[kQ];
    const uint32_t fresh_n = pb.counters[kFresh];
    const uint32_t sbase = pb.counters[kBase];
    const PathState st = pb.cur;
    for (uint32_t base = blockIdx.x * kShadeBlock; base < fresh_n; base += gridDim.x * kShadeBlock) {
        const uint32_t i = base + threadIdx.x;
        const bool fresh = i < fresh_n;
        const uint32_t sidx = sbase + i;
        bool splat = false;
        float4 so = make_float4(0, 0, 0, 0), sd = so, sc = so;
        float4 ro = so, rdir = so, thr = so;
        uint4 meta = make_uint4(0, 0, 0, 0);
        uint32_t job = 0;
        if (fresh) {
            uint32_t local = sidx, wh = a.wh, first = a.first_sample, seed = a.seed;
            int use_mis = a.use_mis;
            if (BATCH) {
                job = find_job(jobs, n_jobs, sidx);
                const BatchJob &J = jobs[job];
                local = sidx - J.offset;
                wh = J.wh;
                first = J.first_sample;
                seed = J.seed;
                use_mis = J.use_mis;
            }
            const DScene &s = BATCH ? jobs[job].s : s1;
            const DCamera &cam = s.cam;
            const uint32_t pixel = local % wh;
            const uint32_t sample = first + local / wh;
            const uint32_t px = pixel % uint32_t(cam.W), py = pixel / uint32_t(cam.W);
            const uint32_t key = path_key(seed, pixel, sample);
            // camera vertex (sampleWe0ByPdf)
            const float3 lens = lens_point(cam, key);
            const float3 cpos = cam_to_world(cam, lens) + f3(cam.ox, cam.oy, cam.oz);
            const float3 cz = f3(cam.zx, cam.zy, cam.zz);
            // t = 1 connection
            const LightSample l = sample_light<ENV>(s, u01(key, kDimLight), u01(key, kDimLight + 1), key);
            if (ENV && a.nee && l.env) {
                // an envmap vertex: the camera looks along its direction; G = |cos| at
                // the camera only, Le1 = 1, visibility to infinity (path.h:64-76, scene.h:98-101)
                const float3 d = l.pos;
                const float G = fabsf(dot3(cz, d));
                float rx = 0.f, ry = 0.f;
                const float we1 = G != 0.f ? camera_we1(cam, lens, d, rx, ry) : 0.f;
                if (we1 != 0.f) {
                    float mis = 1.f;
                    if (use_mis) {
                        const float c = cam_to_local(cam, d).z;
                        const float camPdfA = cam.per_pixel_norm / (c * c * c);
                        mis = l.pdfA / (l.pdfA + camPdfA);
                    }
                    const float3 c = l.contrib * (mis * we1 * G);
                    if (!is_black(c)) {
                        const uint32_t target = kToSplat | (uint32_t(ry) * uint32_t(cam.W) + uint32_t(rx));
                        so = make_float4(cpos.x, cpos.y, cpos.z, INFINITY);
                        sd = make_float4(d.x, d.y, d.z, __uint_as_float(target));
                        sc = make_float4(c.x, c.y, c.z, __uint_as_float(job));
                        splat = true;
                    }
                }
            } else if (a.nee || a.light) {
                const float3 uv = cpos - l.pos;
                const float len2 = dot3(uv, uv);
                const float len = sqrtf(len2);
                const float3 nuv = uv * (1.0f / len);
                const float cos_l = dot3(l.n, nuv);
                const float G = fabsf(cos_l) * fabsf(dot3(cz, nuv)) / len2;
                if (G != 0.f && cos_l >= 0.f) {
                    float rx, ry;
                    const float3 wcam = -nuv;
                    const float we1 = camera_we1(cam, lens, wcam, rx, ry);
                    if (we1 != 0.f) {
                        float mis = 1.f;
                        if (use_mis) {
                            const float c = cam_to_local(cam, wcam).z;
                            float camPdfA = cam.per_pixel_norm / (c * c * c) * fabsf(cos_l) / len2;
                            if (s.med.on) camPdfA *= pass_pdf(s.med, len);
                            mis = l.pdfA / (l.pdfA + camPdfA);
                        }
                        const float k = mis * we1 * G * kInvPi;
                        float3 c = l.contrib * k;
                        if (s.med.on) c = c * transmittance(s.med, len);  // attenuated connection
                        const uint32_t target = kToSplat | (uint32_t(ry) * uint32_t(cam.W) + uint32_t(rx));
                        so = make_float4(l.pos.x, l.pos.y, l.pos.z, len - kRayEps);
                        sd = make_float4(nuv.x, nuv.y, nuv.z, __uint_as_float(target));
                        sc = make_float4(c.x, c.y, c.z, __uint_as_float(job));
                        splat = true;
                    }
                }
            }
            if (a.light) {
                // the light subpath's first ray (pdf not needed: no MIS)
                const float3 dir =
                    to_world(make_frame(l.n), cosine_hemisphere(u01(key, kDimPixel), u01(key, kDimPixel + 1)));
                ro = make_float4(l.pos.x, l.pos.y, l.pos.z, 0.f);
                rdir = make_float4(dir.x, dir.y, dir.z, 0.f);
                thr = make_float4(l.contrib.x, l.contrib.y, l.contrib.z, 0.f);
            } else {
                // primary ray (samplePerPixelWe1CosByPdf)
                float pdfW;
                const float3 dir = primary_dir(cam, key, px, py, lens, pdfW);
                ro = make_float4(cpos.x, cpos.y, cpos.z, 0.f);
                rdir = make_float4(dir.x, dir.y, dir.z, 0.f);
                thr = make_float4(1.f, 1.f, 1.f, use_mis ? pdfW : 0.f);
            }
            meta = make_uint4(key, sidx, job << kBounceBits, 0u);
        }
        if (fresh) {
            const uint32_t e = head + i;
            st.ray_o[e] = ro;
            st.ray_d[e] = rdir;
            st.thr[e] = thr;
            st.rad[e] = make_float4(0.f, 0.f, 0.f, 0.f);
            st.meta[e] = meta;
        }
        // t = 1 connections join the previous shade kernel's shadow rays
        const uint32_t si = block_append(&pb.counters[pb.next.shadow_count], splat, agg);
        if (splat) {
            pb.sh_o[si] = so;
            pb.sh_d[si] = sd;
            pb.sh_c[si] = sc;
        }
    }
}

// ============================================================== extend
// Persistent traversal: a wave claims `chunk` consecutive entries per atomic
// and refills its lanes from that local range as soon as `refill` of them are
// idle, so a wave never waits for its slowest ray and atomics stay rare.
// Waves beyond the work exit at once.  Both are runtime parameters
// (TravTune).  With the quantised nodes (traversal bound by L1 requests),
// on the 1.31M-triangle frame chunk x refill 64x16 / 96x16 / 128x16 / 192x16 /
// 256x16 gave 7,312 / 7,377 / 7,387 / 7,418 / 7,442 Mrays/s and 96x8 / 96x24 /
// 96x32 / 96x48 7,355 / 7,326 / 7,205 / 6,911; in the gap room 256x16 7,938,
// 96x16 7,862, 96x32 7,655: a wave now refills once a quarter of its lanes
// are idle.  wr_autotune searches them per workload.
#ifndef WR_REFILL_LANES
#define WR_REFILL_LANES 16
#endif
#ifndef WR_FETCH_CHUNK
#define WR_FETCH_CHUNK 256
#endif
struct TravTune {
    uint32_t chunk = WR_FETCH_CHUNK;
    uint32_t refill = WR_REFILL_LANES;
};

// The traversal kernels run 8 waves per SIMD without spills: the plain one
// (k_trace_bounce_fast) in 59 VGPRs (7 waves measured 0.8% slower on the
// 1.31M-triangle frame, 1.4% in the gap room), the batched and counting
// instances in 62-64 (the batched one at the compiler's 7 waves made the
// 32-scene batch 3.5% slower: 1.627 against 1.570 s).
#ifndef WR_EXTEND_WAVES
#define WR_EXTEND_WAVES 8
#endif
#define WR_EXTEND_ATTR __attribute__((amdgpu_waves_per_eu(WR_EXTEND_WAVES, WR_EXTEND_WAVES)))

struct WaveRange {
    uint32_t lo = 0, hi = 0;  // wave-uniform
    uint32_t tried = 0;       // XCD shares found empty so far (wave-uniform)
};
// Returns the entry for this lane (>= n when it got none).  The queue's n
// entries are split into kXcds contiguous shares, share x = [n x / 8,
// n (x + 1) / 8), each with its own work counter (`xcd_counters` + x lines):
// workgroups are dispatched to the XCDs round-robin, so block b runs on XCD
// b % 8, and its waves claim chunks of that XCD's share first -- queue order
// follows pixel order, so neighbouring rays, and the nodes they read, stay in
// one XCD's L2 -- then of the other shares in turn, so no XCD idles while work
// is left.  A new chunk is claimed only when the local range is empty;
// `exhausted` (wave-uniform) is set when no share has a chunk left.  All lanes
// of the wave must call it.
__device__ __forceinline__ uint32_t refill(uint32_t *xcd_counters, uint32_t n, bool idle, unsigned long long idle_mask,
                                           WaveRange &wr, bool &exhausted, uint32_t chunk) {
    if (wr.lo == wr.hi) {
        for (;;) {
            const uint32_t x = (blockIdx.x + wr.tried) % kXcds;
            const uint32_t lo = uint32_t(uint64_t(n) * x / kXcds), hi = uint32_t(uint64_t(n) * (x + 1) / kXcds);
            uint32_t got = 0;
            if (lane_id() == 0) got = atomicAdd(&xcd_counters[x * kCounterStride], chunk);
            got = __builtin_amdgcn_readfirstlane(got);  // every lane is active here, so this reads lane 0
            if (got < hi - lo) {
                wr.lo = lo + got;
                wr.hi = min(lo + got + chunk, hi);
                break;
            }
            if (++wr.tried == kXcds) {
                exhausted = true;
                return n;
            }
        }
    }
    const uint32_t k = (uint32_t)__popcll(idle_mask);
    const uint32_t avail = wr.hi - wr.lo;
    const uint32_t r = mbcnt64(idle_mask);
    const uint32_t e = (idle && r < avail) ? wr.lo + r : n;
    wr.lo += min(k, avail);
    return e;
}

// Chunk size for a queue of n entries.  Blocks past n / kBlock exit at once,
// so a short queue (the last bounces of a frame) is worked by ceil(n / 64)
// waves: a chunk of 64 then gives every lane one ray to start at once,
// where a larger chunk leaves fewer waves with work and queues the rest of
// each chunk behind their lanes' first rays.  Long queues keep `chunk`.
__device__ __forceinline__ uint32_t spread_chunk(uint32_t n, uint32_t chunk) {
    const uint32_t per_wave = n / gridDim.x + (n % gridDim.x != 0);
    // wave-uniform; kept in a scalar register (the division runs on the VALU)
    return __builtin_amdgcn_readfirstlane(min(chunk, max(uint32_t(kBlock), per_wave)));
}

// Closest hit for the n active paths of the current state.  `st_ent` is the
// block's LDS stack (kLdsStackDepth x kBlock entries).
template <bool COUNT, bool BATCH>
__device__ __forceinline__ void extend_rays(DScene s1, const BatchJob *__restrict__ jobs, PathBuffers pb, uint2 *spill,
                                            TravTune tune, uint32_t n, uint64_t *st_ent) {
    const uint32_t chunk = spread_chunk(n, tune.chunk);
    LaneStack<kBlock> st(st_ent + lane_id(), reinterpret_cast<uint32_t *>(st_ent) + lane_id(), spill);
    TraversalCounters cnt;
    TravState ts;
    uint32_t p = 0;
    WaveRange wr;
    Geo geo{s1.nodes, s1.tris};
    bool active = false, exhausted = false;
    for (;;) {
        const unsigned long long idle = __ballot(!active);
        if (!exhausted && __popcll(idle) >= tune.refill) {
            const uint32_t i = refill(&pb.counters[kFetchExtX], n, !active, idle, wr, exhausted, chunk);
            if (!active && i < n) {
                p = i;
                const float4 o = pb.cur.ray_o[p], d = pb.cur.ray_d[p];
                if (BATCH) {
                    const DScene &sj = jobs[pb.cur.meta[p].z >> kBounceBits].s;
                    geo = Geo{sj.nodes, sj.tris};
                }
                trav_begin(ts, xyz(o), xyz(d), kRayEps, INFINITY);
                st.sp = 0;
                active = true;
            }
        }
        if (__ballot(active) == 0) {
            if (exhausted) break;
            continue;
        }
        const bool fin = BATCH ? trav_round<COUNT, kBlock>(geo, ts, st, cnt, active)
                               : trav_round<COUNT, kBlock>(s1, ts, st, cnt, active);
        if (active && fin) {
            pb.hit[p] = make_uint2(ts.hit, __float_as_uint(ts.t));
            active = false;
        }
    }
    if (COUNT) {
        wave_add_stat(&pb.stats[2], cnt.nodes);
        wave_add_stat(&pb.stats[3], cnt.tris);
    }
}

// ============================================================== shade
// One vertex of randomWalk with the PathNeeRenderer callback
// (generalsubpathsampler.h:34-130, pathnextevent.h:100-149).  Continuing
// paths are compacted into the next state; a finished path stores its
// radiance in its sample's slot of the pass.
template <bool BATCH, bool ENV = false>
__global__ __launch_bounds__(kShadeBlock) void k_shade(DScene s1, const BatchJob *__restrict__ jobs, PathBuffers pb,
                                                       RenderArgs a) {
    __shared__ uint32_t agg[kWavesPerShadeBlock + 1];
    const uint32_t n = active_count(pb.counters);
    const PathState cs = pb.cur, ns = pb.next;
    for (uint32_t base = blockIdx.x * kShadeBlock; base < n; base += gridDim.x * kShadeBlock) {
        const uint32_t i = base + threadIdx.x;
        bool push_shadow = false, cont = false;
        const bool valid = i < n;
        float4 so = make_float4(0, 0, 0, 0), sd = so, sc = so;
        float4 new_o = so, new_d = so, new_thr = so, rad = so;
        uint4 meta = make_uint4(0, 0, 0, 0);
        uint32_t job = 0;
        if (valid) {
            meta = cs.meta[i];
            rad = cs.rad[i];
            job = meta.z >> kBounceBits;
            const DScene &s = BATCH ? jobs[job].s : s1;
            const int N = BATCH ? jobs[job].N : a.N;
            const int use_mis = BATCH ? jobs[job].use_mis : a.use_mis;
            const uint32_t bounce = meta.z & kBounceMask;
            const int cNum = int(bounce) + 2;  // camera is vertex 1
            const uint2 h = pb.hit[i];
            if (ENV && h.x == kInvalid && s.env_select > 0.f) {
                // a miss is the envmap vertex (scene.h:82-88): an emitter hit with Le1 = 1,
                // Le0 = the lookup, and area pdfs equal to the solid-angle ones
                const float3 rd = xyz(cs.ray_d[i]);
                const float4 thr = cs.thr[i];
                float mis = a.nee ? 0.f : 1.f;
                if (use_mis) {
                    const float lightPdfA = s.env_select * env_pdf_dir(s, rd);
                    mis = thr.w / (lightPdfA + thr.w);
                }
                const float3 c = xyz(thr) * env_eval(s, rd) * mis;
                rad.x += c.x;
                rad.y += c.y;
                rad.z += c.z;
            }
            if (h.x != kInvalid) {
                const float t = __uint_as_float(h.y);
                const float3 ro = xyz(cs.ray_o[i]), rd = xyz(cs.ray_d[i]);
                const float4 v0 = s.tris[3 * size_t(h.x) + 0];
                const float4 v1 = s.tris[3 * size_t(h.x) + 1];
                const float4 v2 = s.tris[3 * size_t(h.x) + 2];
                const float3 pos = ro + rd * t;
                const float3 ng = normalize3(cross3(xyz(v1) - xyz(v0), xyz(v2) - xyz(v0)));
                const int2 gi = s.geom_info[__float_as_uint(v1.w)];
                const float4 thr = cs.thr[i];
                const float3 T = xyz(thr);
                const uint32_t key = meta.x;
                if (gi.y >= 0) {
                    // emitter hit: MIS against light sampling (pathnextevent.h:128-146);
                    // without NEE it is the only estimate (weight 1, primitivepath.h:52-56)
                    const float cos_l = -dot3(ng, rd);
                    if (cos_l >= 0.f) {
                        const DLight &L = s.lights[gi.y];
                        float mis = a.nee ? 0.f : 1.f;
                        if (use_mis) {
                            const float cPdfA = thr.w * fabsf(cos_l) / (t * t);
                            const float lightPdfA = L.pdf_select * L.inv_area;
                            mis = cPdfA / (lightPdfA + cPdfA);
                        }
                        rad.x += mis * T.x * L.L[0];
                        rad.y += mis * T.y * L.L[1];
                        rad.z += mis * T.z * L.L[2];
                    }
                } else if (cNum != N) {
                    const DMat m = s.mats[gi.x];
                    const Frame f = make_frame(ng);
                    const float3 wi = to_local(f, -rd);
                    // NEE with MIS (pathnextevent.h:106-127)
                    const LightSample l = sample_light<ENV>(s, u01(key, kDimLight), u01(key, kDimLight + 1), key);
                    if (ENV && a.nee && l.env) {
                        // toward an envmap vertex: G = |cos| at the surface, visibility to infinity
                        const float3 nuv = l.pos;
                        const float G = fabsf(dot3(ng, nuv));
                        const float3 wo = to_local(f, nuv);
                        const float3 fs = G != 0.f ? bsdf_eval(m, wi, wo) : f3(0.f, 0.f, 0.f);
                        if (!is_black(fs)) {
                            const float mis = use_mis ? l.pdfA / (l.pdfA + bsdf_pdf(m, wi, wo)) : 1.f;
                            const float3 c = T * fs * l.contrib * (mis * G);
                            if (!is_black(c)) {
                                so = make_float4(pos.x, pos.y, pos.z, INFINITY);
                                sd = make_float4(nuv.x, nuv.y, nuv.z, 0.f);
                                sc = make_float4(c.x, c.y, c.z, 0.f);
                                push_shadow = true;
                            }
                        }
                    }
                    const float3 uv = l.pos - pos;
                    const float len2 = dot3(uv, uv);
                    const float len = sqrtf(len2);
                    const float3 nuv = uv * (1.0f / len);
                    const float cos_lv = -dot3(l.n, nuv);
                    const float G = fabsf(dot3(ng, nuv)) * fabsf(cos_lv) / len2;
                    if (a.nee && !(ENV && l.env) && G != 0.f && cos_lv >= 0.f) {
                        const float3 wo = to_local(f, nuv);
                        const float3 fs = bsdf_eval(m, wi, wo);
                        if (!is_black(fs)) {
                            float mis = 1.f;
                            if (use_mis) {
                                const float bsdfPdfA = bsdf_pdf(m, wi, wo) * fabsf(cos_lv) / len2;
                                mis = l.pdfA / (l.pdfA + bsdfPdfA);
                            }
                            const float k = mis * G * kInvPi;
                            const float3 c = T * fs * l.contrib * k;
                            if (!is_black(c)) {
                                so = make_float4(pos.x, pos.y, pos.z, len - kRayEps);
                                sd = make_float4(nuv.x, nuv.y, nuv.z, 0.f);
                                sc = make_float4(c.x, c.y, c.z, 0.f);
                                push_shadow = true;
                            }
                        }
                    }
                    // BSDF sampling (generalsubpathsampler.h:81-130); here cNum < N
                    const uint32_t dim = kDimBounce0 + kDimsPerBounce * bounce + kBdBsdf;
                    float3 wo;
                    float pdf = 0.f;
                    const float3 w = bsdf_sample(m, wi, u01(key, dim), u01(key, dim + 1), wo, pdf);
                    if (!is_black(w)) {
                        const float3 nd = to_world(f, wo);
                        new_o = make_float4(pos.x, pos.y, pos.z, 0.f);
                        new_d = make_float4(nd.x, nd.y, nd.z, 0.f);
                        new_thr = make_float4(w.x * T.x, w.y * T.y, w.z * T.z, use_mis ? pdf : 0.f);
                        cont = true;
                    }
                }
            }
        }
        const uint32_t qi = block_append(&pb.counters[kQn], cont, agg);
        if (cont) {
            ns.ray_o[qi] = new_o;
            ns.ray_d[qi] = new_d;
            ns.thr[qi] = new_thr;
            ns.rad[qi] = rad;
            ns.meta[qi] = make_uint4(meta.x, meta.y, meta.z + 1, 0u);
        } else if (valid) {
            pb.sample_rad[meta.y] = rad;
        }
        const uint32_t target = cont ? (kToNextPath | qi) : (kToSample | meta.y);
        sd.w = __uint_as_float(target);
        sc.w = __uint_as_float(job);
        const uint32_t si = block_append(&pb.counters[pb.cur.shadow_count], push_shadow, agg);
        if (push_shadow) {
            pb.sh_o[si] = so;
            pb.sh_d[si] = sd;
            pb.sh_c[si] = sc;
        }
    }
}

// ============================================================== light subpaths
// One vertex of a light subpath (LightSplattingRenderer, lightsplatting.h:91-105;
// randomWalk, generalsubpathsampler.h:34-160).  Vertex k of the subpath (the
// light vertex is 1; the first hit is bounce 0, vertex 2) exists for k <= N - 1.
// A surface vertex connects to the camera: G We1 f T with f evaluated from the
// previous vertex toward the camera, splatted at the camera raster through a
// shadow ray; then, before vertex N - 1, the BSDF samples the next direction.
// An emitter hit ends the subpath.  No medium (the reference's sampler here
// is not volumetric).
__global__ __launch_bounds__(kShadeBlock) void k_shade_light(DScene s, PathBuffers pb, RenderArgs a) {
    __shared__ uint32_t agg[kWavesPerShadeBlock + 1];
    const uint32_t n = active_count(pb.counters);
    const PathState cs = pb.cur, ns = pb.next;
    const DCamera &cam = s.cam;
    for (uint32_t base = blockIdx.x * kShadeBlock; base < n; base += gridDim.x * kShadeBlock) {
        const uint32_t i = base + threadIdx.x;
        bool push_shadow = false, cont = false;
        const bool valid = i < n;
        float4 so = make_float4(0, 0, 0, 0), sd = so, sc = so;
        float4 new_o = so, new_d = so, new_thr = so, rad = so;
        uint4 meta = make_uint4(0, 0, 0, 0);
        if (valid) {
            meta = cs.meta[i];
            rad = cs.rad[i];
            const uint32_t bounce = meta.z & kBounceMask;
            const int k = int(bounce) + 2;  // vertex of the light subpath
            const uint2 h = pb.hit[i];
            if (h.x != kInvalid && k <= a.N - 1) {
                const float t = __uint_as_float(h.y);
                const float3 ro = xyz(cs.ray_o[i]), rd = xyz(cs.ray_d[i]);
                const float4 v0 = s.tris[3 * size_t(h.x) + 0];
                const float4 v1 = s.tris[3 * size_t(h.x) + 1];
                const float4 v2 = s.tris[3 * size_t(h.x) + 2];
                const int2 gi = s.geom_info[__float_as_uint(v1.w)];
                if (gi.y < 0) {
                    const float3 pos = ro + rd * t;
                    const float3 ng = normalize3(cross3(xyz(v1) - xyz(v0), xyz(v2) - xyz(v0)));
                    const float3 T = xyz(cs.thr[i]);
                    const uint32_t key = meta.x;
                    const DMat m = s.mats[gi.x];
                    const Frame f = make_frame(ng);
                    const float3 wi = to_local(f, -rd);
                    // connection to the camera vertex
                    const float3 lens = lens_point(cam, key);
                    const float3 cpos = cam_to_world(cam, lens) + f3(cam.ox, cam.oy, cam.oz);
                    const float3 uv = cpos - pos;
                    const float len2 = dot3(uv, uv);
                    const float len = sqrtf(len2);
                    const float3 nuv = uv * (1.0f / len);
                    const float G = fabsf(dot3(ng, nuv)) * fabsf(dot3(f3(cam.zx, cam.zy, cam.zz), nuv)) / len2;
                    float rx = 0.f, ry = 0.f;
                    const float we1 = G != 0.f ? camera_we1(cam, lens, -nuv, rx, ry) : 0.f;
                    if (we1 != 0.f) {
                        const float3 fs = bsdf_eval(m, wi, to_local(f, nuv));
                        const float3 c = T * fs * (G * we1);
                        if (!is_black(c)) {
                            const uint32_t target = kToSplat | (uint32_t(ry) * uint32_t(cam.W) + uint32_t(rx));
                            so = make_float4(pos.x, pos.y, pos.z, len - kRayEps);
                            sd = make_float4(nuv.x, nuv.y, nuv.z, __uint_as_float(target));
                            sc = make_float4(c.x, c.y, c.z, 0.f);
                            push_shadow = true;
                        }
                    }
                    if (k < a.N - 1) {
                        const uint32_t dim = kDimBounce0 + kDimsPerBounce * bounce + kBdBsdf;
                        float3 wo;
                        float pdf = 0.f;
                        const float3 w = bsdf_sample(m, wi, u01(key, dim), u01(key, dim + 1), wo, pdf);
                        if (!is_black(w)) {
                            const float3 nd = to_world(f, wo);
                            new_o = make_float4(pos.x, pos.y, pos.z, 0.f);
                            new_d = make_float4(nd.x, nd.y, nd.z, 0.f);
                            new_thr = make_float4(w.x * T.x, w.y * T.y, w.z * T.z, 0.f);
                            cont = true;
                        }
                    }
                }
            }
        }
        const uint32_t qi = block_append(&pb.counters[kQn], cont, agg);
        if (cont) {
            ns.ray_o[qi] = new_o;
            ns.ray_d[qi] = new_d;
            ns.thr[qi] = new_thr;
            ns.rad[qi] = rad;
            ns.meta[qi] = make_uint4(meta.x, meta.y, meta.z + 1, 0u);
        } else if (valid) {
            pb.sample_rad[meta.y] = rad;
        }
        const uint32_t si = block_append(&pb.counters[pb.cur.shadow_count], push_shadow, agg);
        if (push_shadow) {
            pb.sh_o[si] = so;
            pb.sh_d[si] = sd;
            pb.sh_c[si] = sc;
        }
    }
}

// ============================================================== extended shade
// k_shade with learned guiding and/or a homogeneous medium.
//
// Guiding: the direction at a vertex is drawn by one-sample MIS from the
// network's lobe mixture (probability alpha) or the BSDF / phase function, and
// weighted by f / (alpha q + (1 - alpha) p); the same combined density enters
// the NEE and emitter-hit MIS weights, so the estimator stays unbiased for any
// network.  k_guide_infer draws the direction and evaluates q.  Selected paths
// append a training record at every scattering vertex; the record is closed
// when the path's next vertex is shaded (radiance snapshot).
//
// Medium (homogeneous.h:11-66, generalsubpathsampler.h:40-53,131-154): each
// segment first samples a free flight on one colour channel; a flight that ends
// before the surface makes a medium vertex (throughput sigma_s Tr / pdf, NEE
// through the HG phase function, phase sampling), otherwise the surface is
// reached with throughput Tr / P(pass).  Shadow contributions carry the
// transmittance of their segment, and the MIS densities the pass probability.
__device__ __forceinline__ float guide_alpha(const DMat &m, float frac) {
    // sharp Phong lobes are sampled well by the BSDF itself
    return m.type == kPhong ? frac * fminf(1.f, fmaxf(0.1f, 30.f / m.n)) : frac;
}

// Network inference for the current state, for every vertex that k_shade_ext
// samples a direction at -- a surface that is not an emitter or the last
// vertex, or (with a medium) a scattering event before the last vertex -- one
// forward pass per wave of 64 entries on MFMA.  With the lobes on chip it also
// evaluates what shading needs of them: it draws the vertex's direction (the
// network with probability alpha, else the BSDF or phase function, from the
// same random numbers k_shade would use) and returns q at that direction and
// at the direction of the vertex's light sample.  The free-flight decision
// repeats k_shade_ext's.
// Network outputs reach their record's lane through v_permlane32_swap, so LDS
// holds only the weights (37 KB BF16, 18 KB e4m3) and the kernel runs 3 waves
// per SIMD (155-166 VGPRs): 2/3/4 waves took 91.7/82.4/81.4 ms of guided
// shading per 32 spp at 1080p (4 spills registers), 88.7 with the former
// 42 KB LDS stage at 2 waves.
#ifndef WR_INFER_WAVES
#define WR_INFER_WAVES 3
#endif
#ifndef WR_INFER_LDS_WEIGHTS
#define WR_INFER_LDS_WEIGHTS 1
#endif
constexpr int kInferBlock = 256;  // 4 waves share one LDS copy of the weights
// FP8: the e4m3 variant of the network (guide::forward8).
template <bool MEDIUM, bool FP8>
__global__ __launch_bounds__(kInferBlock) __attribute__((amdgpu_waves_per_eu(WR_INFER_WAVES, WR_INFER_WAVES))) void k_guide_infer(
    DScene s, PathState cs, const uint2 *__restrict__ hit, const uint32_t *__restrict__ counters, int N, guide::Net gnet,
    float frac, uint32_t max_bounce, guide::Draw out) {
    using OpT = typename std::conditional<FP8, guide::u32x2, guide::u32x4>::type;
#if WR_INFER_LDS_WEIGHTS
    __shared__ OpT w_ops[guide::kFwdOps];
#endif
    __shared__ float w_bias[guide::kBiasPacked];
    __shared__ float w_scale[guide::kLayers];
    const uint32_t n = active_count(counters);
    if (blockIdx.x * kInferBlock >= n) return;
#if WR_INFER_LDS_WEIGHTS
    const OpT *src;
    if constexpr (FP8) src = gnet.fwd8;
    else src = gnet.fwd;
    for (int j = threadIdx.x; j < guide::kFwdOps; j += kInferBlock) w_ops[j] = src[j];
#endif
    for (int j = threadIdx.x; j < guide::kBiasPacked; j += kInferBlock) w_bias[j] = gnet.bias[j];
    if (FP8 && threadIdx.x < guide::kLayers) w_scale[threadIdx.x] = gnet.wscale[threadIdx.x];
    __syncthreads();
    guide::Net net = gnet;
    if constexpr (FP8) {
#if WR_INFER_LDS_WEIGHTS
        net.fwd8 = w_ops;
#endif
        net.wscale = w_scale;
    } else {
#if WR_INFER_LDS_WEIGHTS
        net.fwd = w_ops;
#endif
    }
    net.bias = w_bias;
    for (uint32_t base = blockIdx.x * kInferBlock; base < n; base += gridDim.x * kInferBlock) {
        const uint32_t i = base + threadIdx.x;
        bool need = false, med = false;
        float3 pos = f3(0.f, 0.f, 0.f), ng = f3(0.f, 1.f, 0.f), nf = ng, rd = f3(0.f, 0.f, 1.f);
        int mat = 0;
        uint4 meta = make_uint4(0, 0, 0, 0);
        if (i < n) {
            const uint2 h = hit[i];
            meta = cs.meta[i];
            if (int(meta.z) + 2 != N && meta.z < max_bounce) {
                float tf = 0.f;
                med = MEDIUM && free_flight(s.med, meta.x, meta.z, h.x != kInvalid ? __uint_as_float(h.y) : INFINITY, tf);
                if (med) {
                    need = true;
                    rd = xyz(cs.ray_d[i]);
                    pos = xyz(cs.ray_o[i]) + rd * tf;
                    nf = -rd;  // no surface: the incoming direction stands in for the normal
                } else if (h.x != kInvalid) {
                    const float4 v0 = s.tris[3 * size_t(h.x) + 0];
                    const float4 v1 = s.tris[3 * size_t(h.x) + 1];
                    const float4 v2 = s.tris[3 * size_t(h.x) + 2];
                    const int2 gi = s.geom_info[__float_as_uint(v1.w)];
                    if (gi.y < 0) {
                        need = true;
                        mat = gi.x;
                        rd = xyz(cs.ray_d[i]);
                        pos = xyz(cs.ray_o[i]) + rd * __uint_as_float(h.y);
                        ng = normalize3(cross3(xyz(v1) - xyz(v0), xyz(v2) - xyz(v0)));
                        nf = dot3(ng, rd) > 0.f ? -ng : ng;
                    }
                }
            }
        }
        if (__ballot(need)) {
            float L[guide::kOut];  // this lane's record's raw outputs, then its decoded lobes
            if constexpr (FP8) guide::forward8(net, guide::encode(pos, nf, -rd, net.box), L);
            else guide::forward(net, guide::encode(pos, nf, -rd, net.box), L);
            if (need) {
                float wl[guide::kLobes];
                guide::decode(L, wl);
                const uint32_t key = meta.x;
                const LightSample l = sample_light(s, u01(key, kDimLight), u01(key, kDimLight + 1), key);
                const float3 dl = normalize3(l.pos - pos);
                const uint32_t dim = kDimBounce0 + kDimsPerBounce * meta.z + kBdBsdf;
                const float u0 = u01(key, dim), u1 = u01(key, dim + 1), us = u01(key, dim + 2);
                float alpha = frac;
                DMat m{};
                if (!med) {
                    m = s.mats[mat];
                    alpha = guide_alpha(m, frac);
                }
                // At a surface the mixture is folded onto the hemisphere of nf (the
                // BSDFs reflect only): a draw below the surface is mirrored across
                // the tangent plane, so the density is q(w) + q(R w) above and 0
                // below -- normalised, and no network draw is lost to f = 0.
                auto q = [&](float3 w) {
                    if (med) return guide::pdf(L, w);
                    const float c = dot3(w, nf);
                    return c < 0.f ? 0.f : guide::pdf(L, w) + guide::pdf(L, w - nf * (2.f * c));
                };
                float3 d;
                if (us < alpha) {
                    d = guide::sample(L, wl, us * guide::frcp(alpha), u0, u1);
                    const float c = dot3(d, nf);
                    if (!med && c < 0.f) d = d - nf * (2.f * c);
                } else if (med) {
                    d = hg_sample(s.med.g, rd, u0, u1);
                } else {
                    const Frame f = make_frame(ng);
                    float3 wo = f3(0.f, 0.f, 0.f);
                    float pb = 0.f;
                    bsdf_sample(m, to_local(f, -rd), u0, u1, wo, pb);
                    d = normalize3(to_world(f, wo));
                }
                out.dir[i] = make_float4(d.x, d.y, d.z, q(d));
                out.q_nee[i] = q(dl);
            }
        }
    }
}

template <bool GUIDED, bool MEDIUM>
__global__ __launch_bounds__(kShadeBlock) void k_shade_ext(DScene s, PathBuffers pb, RenderArgs a, guide::Shade g) {
    __shared__ uint32_t agg[kWavesPerShadeBlock + 1];
    const uint32_t n = active_count(pb.counters);
    const PathState cs = pb.cur, ns = pb.next;
    const float3 zero = f3(0.f, 0.f, 0.f);
    for (uint32_t base = blockIdx.x * kShadeBlock; base < n; base += gridDim.x * kShadeBlock) {
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < n;
        // ---- the vertex
        uint4 meta = make_uint4(0, 0, 0, 0);
        uint2 h = make_uint2(kInvalid, 0u);
        float3 pos = zero, ng = f3(0.f, 1.f, 0.f), rd = f3(0.f, 0.f, 1.f);
        int2 gi = make_int2(0, -1);
        bool in_medium = false;
        float seg_t = 0.f;
        if (valid) {
            meta = cs.meta[i];
            h = pb.hit[i];
            rd = xyz(cs.ray_d[i]);
            seg_t = h.x != kInvalid ? __uint_as_float(h.y) : INFINITY;
            if (MEDIUM) {
                float tf = 0.f;
                in_medium = free_flight(s.med, meta.x, meta.z, seg_t, tf);
                if (in_medium) seg_t = tf;
            }
            if (in_medium) {
                pos = xyz(cs.ray_o[i]) + rd * seg_t;
            } else if (h.x != kInvalid) {
                pos = xyz(cs.ray_o[i]) + rd * seg_t;
                const float4 v0 = s.tris[3 * size_t(h.x) + 0];
                const float4 v1 = s.tris[3 * size_t(h.x) + 1];
                const float4 v2 = s.tris[3 * size_t(h.x) + 2];
                ng = normalize3(cross3(xyz(v1) - xyz(v0), xyz(v2) - xyz(v0)));
                gi = s.geom_info[__float_as_uint(v1.w)];
            }
        }
        const bool hit_vertex = valid && !in_medium && h.x != kInvalid;
        const bool last = int(meta.z) + 2 == a.N;
        const bool surf = hit_vertex && gi.y < 0 && !last;
        const bool med = valid && in_medium && !last;
        const float3 nf = in_medium ? -rd : (dot3(ng, rd) > 0.f ? -ng : ng);  // faces the incoming ray
        const bool need = GUIDED && (surf || med) && g.alpha > 0.f && meta.z < g.max_bounce;
        float4 gdir = make_float4(0.f, 0.f, 0.f, 0.f);  // k_guide_infer's direction and q there
        float q_nee = 0.f;
        if (GUIDED && need) {
            gdir = g.draw.dir[i];
            q_nee = g.draw.q_nee[i];
        }
        // ---- shading
        bool push_shadow = false, cont = false;
        float4 so = make_float4(0, 0, 0, 0), sd = so, sc = so;
        float4 new_o = so, new_d = so, new_thr = so, rad = so;
        float3 wdir = zero, T = zero;
        const uint32_t key = meta.x, bounce = meta.z;
        float prev_pdfw = 0.f;
        if (valid) {
            rad = cs.rad[i];
            // radiance before this vertex closes the record of the previous one
            if (GUIDED && meta.w) g.rec.rad[meta.w - 1] = rad;
            const float4 thr = cs.thr[i];
            prev_pdfw = thr.w;
            T = xyz(thr);
            if (MEDIUM && (in_medium || h.x != kInvalid)) {
                // free-flight weight of the segment
                const float3 tr = transmittance(s.med, seg_t);
                if (in_medium) T = T * f3(s.med.sig_s[0], s.med.sig_s[1], s.med.sig_s[2]) * tr * (1.f / scatter_pdf(s.med, seg_t));
                else T = T * tr * (1.f / pass_pdf(s.med, seg_t));
            }
        }
        if (hit_vertex && gi.y >= 0) {
            // emitter hit: MIS against light sampling (pathnextevent.h:128-146)
            const float cos_l = -dot3(ng, rd);
            if (cos_l >= 0.f) {
                const DLight &L = s.lights[gi.y];
                float mis = a.nee ? 0.f : 1.f;
                if (a.use_mis) {
                    float cPdfA = prev_pdfw * fabsf(cos_l) / (seg_t * seg_t);
                    if (MEDIUM) cPdfA *= pass_pdf(s.med, seg_t);
                    const float lightPdfA = L.pdf_select * L.inv_area;
                    mis = cPdfA / (lightPdfA + cPdfA);
                }
                rad.x += mis * T.x * L.L[0];
                rad.y += mis * T.y * L.L[1];
                rad.z += mis * T.z * L.L[2];
            }
        }
        if (surf || med) {
            DMat m{};
            Frame f{};
            float3 wi = zero;
            if (surf) {
                m = s.mats[gi.x];
                f = make_frame(ng);
                wi = to_local(f, -rd);
            }
            const float alpha = need ? (surf ? guide_alpha(m, g.alpha) : g.alpha) : 0.f;
            // NEE with MIS against the density that samples directions here (pathnextevent.h:106-127)
            const LightSample l = sample_light(s, u01(key, kDimLight), u01(key, kDimLight + 1), key);
            const float3 uv = l.pos - pos;
            const float len2 = dot3(uv, uv);
            const float len = sqrtf(len2);
            const float3 nuv = uv * (1.0f / len);
            const float cos_lv = -dot3(l.n, nuv);
            const float G = (med ? 1.f : fabsf(dot3(ng, nuv))) * fabsf(cos_lv) / len2;
            if (a.nee && G != 0.f && cos_lv >= 0.f) {
                float3 fs;
                float p;
                if (med) {
                    const float ph = hg_pdf(s.med.g, dot3(rd, nuv));
                    fs = f3(ph, ph, ph);
                    p = ph;
                } else {
                    const float3 wo = to_local(f, nuv);
                    fs = bsdf_eval(m, wi, wo);
                    p = bsdf_pdf(m, wi, wo);
                }
                if (!is_black(fs)) {
                    float mis = 1.f;
                    if (a.use_mis) {
                        if (GUIDED && need) p = alpha * q_nee + (1.f - alpha) * p;
                        float dirPdfA = p * fabsf(cos_lv) / len2;
                        if (MEDIUM) dirPdfA *= pass_pdf(s.med, len);
                        mis = l.pdfA / (l.pdfA + dirPdfA);
                    }
                    const float k = mis * G * kInvPi;
                    float3 c = T * fs * l.contrib * k;
                    if (MEDIUM) c = c * transmittance(s.med, len);
                    if (!is_black(c)) {
                        so = make_float4(pos.x, pos.y, pos.z, len - kRayEps);
                        sd = make_float4(nuv.x, nuv.y, nuv.z, 0.f);
                        sc = make_float4(c.x, c.y, c.z, 0.f);
                        push_shadow = true;
                    }
                }
            }
            // direction: BSDF or phase function, one-sample MIS with the network when guided
            const uint32_t dim = kDimBounce0 + kDimsPerBounce * bounce + kBdBsdf;
            const float u0 = u01(key, dim), u1 = u01(key, dim + 1);
            float3 nd = zero, w = zero;
            float pdf = 0.f;
            if (GUIDED && need) {
                // the direction k_guide_infer drew from the mixture alpha q + (1 - alpha) p;
                // whichever component drew it, its weight is f cos / (alpha q + (1 - alpha) p)
                nd = xyz(gdir);
                float3 fv;
                float pd;
                if (med) {
                    pd = hg_pdf(s.med.g, dot3(rd, nd));
                    fv = f3(pd, pd, pd);
                } else {
                    const float3 wo = to_local(f, nd);
                    fv = bsdf_eval(m, wi, wo) * fabsf(wo.y);
                    pd = bsdf_pdf(m, wi, wo);
                }
                pdf = alpha * gdir.w + (1.f - alpha) * pd;
                if (pdf > 0.f && !is_black(fv)) w = fv * (1.f / pdf);
            } else if (med) {
                nd = hg_sample(s.med.g, rd, u0, u1);
                pdf = hg_pdf(s.med.g, dot3(rd, nd));
                w = f3(1.f, 1.f, 1.f);
            } else {
                float3 wo = zero;
                w = bsdf_sample(m, wi, u0, u1, wo, pdf);
                nd = to_world(f, wo);
            }
            if (!is_black(w) && pdf > 0.f) {
                wdir = nd;
                new_o = make_float4(pos.x, pos.y, pos.z, 0.f);
                new_d = make_float4(nd.x, nd.y, nd.z, 0.f);
                new_thr = make_float4(w.x * T.x, w.y * T.y, w.z * T.z, a.use_mis ? pdf : 0.f);
                cont = true;
            }
        }
        uint32_t rec_id = 0;
        if (GUIDED && g.collect) {
            const bool want = cont && bounce < g.max_bounce && float(mix32(key ^ 0x68bc21ebu) >> 8) * 0x1p-24f < g.record_prob;
            const uint32_t r = block_append(g.rec.count, want, agg);
            if (want && r < g.rec.capacity) {
                g.rec.pos[r] = make_float4(pos.x, pos.y, pos.z, __uint_as_float(meta.y));
                g.rec.nrm[r] = make_float4(nf.x, nf.y, nf.z, 0.f);
                g.rec.wo[r] = make_float4(-rd.x, -rd.y, -rd.z, 0.f);
                g.rec.wi[r] = make_float4(wdir.x, wdir.y, wdir.z, 0.f);
                g.rec.thr[r] = make_float4(T.x, T.y, T.z, 0.f);
                g.rec.rad[r] = make_float4(0.f, 0.f, 0.f, 0.f);
                rec_id = r + 1;
            }
        }
        const uint32_t qi = block_append(&pb.counters[kQn], cont, agg);
        if (cont) {
            ns.ray_o[qi] = new_o;
            ns.ray_d[qi] = new_d;
            ns.thr[qi] = new_thr;
            ns.rad[qi] = rad;
            ns.meta[qi] = make_uint4(meta.x, meta.y, meta.z + 1, rec_id);
        } else if (valid) {
            pb.sample_rad[meta.y] = rad;
        }
        const uint32_t target = cont ? (kToNextPath | qi) : (kToSample | meta.y);
        sd.w = __uint_as_float(target);
        const uint32_t si = block_append(&pb.counters[pb.cur.shadow_count], push_shadow, agg);
        if (push_shadow) {
            pb.sh_o[si] = so;
            pb.sh_d[si] = sd;
            pb.sh_c[si] = sc;
        }
    }
}

// ============================================================== shadow
// Any hit for the n entries of the shadow queue the previous shade kernel (and
// this iteration's regeneration) filled; a visible connection adds its
// contribution to the radiance of its path (continuing: by then in the
// current state; finished: its sample), or to the FP64 splat film for
// light->camera (t = 1) connections.  `st_node` is the block's LDS stack of
// node refs (any hit keeps no entry distances).
template <bool COUNT, bool BATCH>
__device__ __forceinline__ void shadow_rays(DScene s1, const BatchJob *__restrict__ jobs, PathBuffers pb, uint2 *spill,
                                            TravTune tune, uint32_t n, uint32_t *st_node) {
    const uint32_t chunk = spread_chunk(n, tune.chunk);
    LaneStack<kBlock> st(nullptr, st_node + lane_id(), spill);
    TraversalCounters cnt;
    TravState ts;
    uint32_t i = 0;
    WaveRange wr;
    Geo geo{s1.nodes, s1.tris};
    bool active = false, exhausted = false;
    for (;;) {
        const unsigned long long idle = __ballot(!active);
        if (!exhausted && __popcll(idle) >= tune.refill) {
            const uint32_t j = refill(&pb.counters[kFetchShdX], n, !active, idle, wr, exhausted, chunk);
            if (!active && j < n) {
                i = j;
                const float4 o = pb.sh_o[i], d = pb.sh_d[i];
                if (BATCH) {
                    const DScene &sj = jobs[__float_as_uint(pb.sh_c[i].w)].s;
                    geo = Geo{sj.nodes, sj.tris};
                }
                trav_begin(ts, xyz(o), xyz(d), kRayEps, o.w);
                st.sp = 0;
                active = true;
            }
        }
        if (__ballot(active) == 0) {
            if (exhausted) break;
            continue;
        }
        const bool fin = BATCH ? trav_round_any<COUNT, kBlock>(geo, ts, st, cnt, active)
                               : trav_round_any<COUNT, kBlock>(s1, ts, st, cnt, active);
        if (active && fin) {
            active = false;
            if (ts.hit == kInvalid) {
                const float4 c = pb.sh_c[i];
                const uint32_t target = __float_as_uint(pb.sh_d[i].w);
                const uint32_t idx = target & kTargetMask;
                if ((target & ~kTargetMask) == kToSplat) {
                    double *dst = (BATCH ? jobs[__float_as_uint(c.w)].film_splat : pb.film_splat) + 3 * size_t(idx);
                    atomicAdd(dst + 0, double(c.x));
                    atomicAdd(dst + 1, double(c.y));
                    atomicAdd(dst + 2, double(c.z));
                } else {
                    float4 *r = (target & ~kTargetMask) == kToSample ? pb.sample_rad + idx : pb.cur.rad + idx;
                    float4 rr = *r;
                    rr.x += c.x; rr.y += c.y; rr.z += c.z;
                    *r = rr;
                }
            }
        }
    }
    if (COUNT) {
        wave_add_stat(&pb.stats[4], cnt.nodes);
        wave_add_stat(&pb.stats[5], cnt.tris);
    }
}

// ============================================================== traversal launch
// One launch per iteration traces the closest-hit rays of the current state
// and the shadow queue of the previous shade kernel: the two are independent
// (both read only what that shade kernel and this iteration's regeneration
// wrote), and in one launch each queue's last, slowest rays overlap the other
// queue's work instead of ending a launch of their own.  A block with
// closest-hit work traces it first; the shadow queue is dealt to the blocks
// from the first one without closest-hit work on (block index rotated by the
// number of blocks the closest-hit rays occupy), so short queues run side by
// side and long ones back to back.  Both share the block's LDS stack.
template <bool COUNT, bool BATCH>
__device__ __forceinline__ void trace_bounce(DScene s1, const BatchJob *__restrict__ jobs, PathBuffers pb, uint2 *spill,
                                             TravTune tune) {
    __shared__ uint64_t st_ent[kLdsStackDepth * kBlock];
    const uint32_t n_ext = active_count(pb.counters);
    const uint32_t n_shd = pb.counters[pb.next.shadow_count];
    const uint32_t g = gridDim.x;
    const uint32_t ext_blocks = min(g, n_ext / kBlock + (n_ext % kBlock != 0));
    if (blockIdx.x < ext_blocks) extend_rays<COUNT, BATCH>(s1, jobs, pb, spill, tune, n_ext, st_ent);
    const uint32_t sb = (blockIdx.x + g - ext_blocks) % g;
    if (sb * kBlock < n_shd)
        shadow_rays<COUNT, BATCH>(s1, jobs, pb, spill, tune, n_shd, reinterpret_cast<uint32_t *>(st_ent));
}
template <bool COUNT, bool BATCH>
__global__ __launch_bounds__(kBlock) WR_EXTEND_ATTR void k_trace_bounce(DScene s1, const BatchJob *__restrict__ jobs,
                                                                        PathBuffers pb, uint2 *spill, TravTune tune) {
    trace_bounce<COUNT, BATCH>(s1, jobs, pb, spill, tune);
}
// The plain kernel of wr_render_nee / wr_render_guided and the other
// integrators' loops.
__global__ __launch_bounds__(kBlock) WR_EXTEND_ATTR void k_trace_bounce_fast(DScene s1, PathBuffers pb, uint2 *spill,
                                                                             TravTune tune) {
    trace_bounce<false, false>(s1, nullptr, pb, spill, tune);
}

// ============================================================== accumulate
// Film::addSample for a pass (film.h:42-48): the samples of pixel p are
// p, p + wh, p + 2 wh, ...; their FP32 radiance is summed in FP64, in sample
// order, so the film is bit-reproducible.
__global__ void k_accumulate(const float4 *__restrict__ sample_rad, double *film_base, uint32_t wh, uint32_t spp) {
    for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < wh; px += gridDim.x * blockDim.x) {
        double r = 0.0, g = 0.0, b = 0.0;
        for (uint32_t k = 0; k < spp; k++) {
            const float4 v = sample_rad[size_t(k) * wh + px];
            r += v.x; g += v.y; b += v.z;
        }
        double *dst = film_base + 3 * size_t(px);
        dst[0] += r; dst[1] += g; dst[2] += b;
    }
}

// ============================================================== trace kernels (parity / benchmark)
template <bool ANY>
__global__ __launch_bounds__(kBlock) void k_trace(DScene s, const float4 *__restrict__ rays, uint32_t n, uint2 *hits,
                                                  uint8_t *occ, uint2 *spill) {
    __shared__ uint64_t st_ent[kLdsStackDepth * kBlock];
    LaneStack<kBlock> st(st_ent + lane_id(), reinterpret_cast<uint32_t *>(st_ent) + lane_id(), spill);
    TraversalCounters cnt;
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        const uint32_t i = base + lane_id();
        if (i < n) {
            const float4 a = rays[2 * size_t(i)], b = rays[2 * size_t(i) + 1];
            const RaySetup r = setup_ray(xyz(a), xyz(b), a.w);
            float t = b.w;
            const uint32_t tri = traverse_rounds<ANY, kBlock>(s, r, t, st, cnt);
            if (ANY) occ[i] = tri != kInvalid;
            else hits[i] = make_uint2(tri, __float_as_uint(tri == kInvalid ? INFINITY : t));
        }
    }
}

// VisualizeGeometryNormal::render (visualizegeomnormal.h:32-74): per sample
// the camera ray's first hit adds (1, 1, 1) on an emitter and the unit
// geometric normal (v1 - v0) x (v2 - v0), unflipped, as RGB on any other
// surface; a miss adds nothing.  The camera sample (lens point, pixel jitter)
// is the path tracers' one.  One lane per pixel sums its samples in sample
// order into the FP64 base film, so the film is bit-reproducible.
__global__ __launch_bounds__(kBlock) void k_visualize(DScene s, uint32_t wh, uint32_t spp, uint32_t first_sample,
                                                      uint32_t seed, double *film_base, uint2 *spill) {
    __shared__ uint64_t st_ent[kLdsStackDepth * kBlock];
    LaneStack<kBlock> st(st_ent + lane_id(), reinterpret_cast<uint32_t *>(st_ent) + lane_id(), spill);
    TraversalCounters cnt;
    const DCamera &cam = s.cam;
    for (uint32_t pixel = blockIdx.x * kBlock + lane_id(); pixel < wh; pixel += gridDim.x * kBlock) {
        const uint32_t px = pixel % uint32_t(cam.W), py = pixel / uint32_t(cam.W);
        double r = 0.0, g = 0.0, b = 0.0;
        for (uint32_t k = 0; k < spp; k++) {
            const uint32_t key = path_key(seed, pixel, first_sample + k);
            const float3 lens = lens_point(cam, key);
            const float3 cpos = cam_to_world(cam, lens) + f3(cam.ox, cam.oy, cam.oz);
            float pdfW;
            const float3 dir = primary_dir(cam, key, px, py, lens, pdfW);
            float t = INFINITY;
            const uint32_t tri = traverse_rounds<false, kBlock>(Geo4{s.fnodes, s.tris}, setup_ray(cpos, dir, kRayEps),
                                                                t, st, cnt);
            if (tri == kInvalid) {
                if (s.env_select > 0.f) { r += 1.0; g += 1.0; b += 1.0; }  // the envmap vertex is a light vertex
                continue;
            }
            const float4 v0 = s.tris[3 * size_t(tri) + 0], v1 = s.tris[3 * size_t(tri) + 1],
                         v2 = s.tris[3 * size_t(tri) + 2];
            if (s.geom_info[__float_as_uint(v1.w)].y >= 0) {
                r += 1.0; g += 1.0; b += 1.0;
            } else {
                const float3 ng = normalize3(cross3(xyz(v1) - xyz(v0), xyz(v2) - xyz(v0)));
                r += ng.x; g += ng.y; b += ng.z;
            }
        }
        double *dst = film_base + 3 * size_t(pixel);
        dst[0] += r; dst[1] += g; dst[2] += b;
    }
}

// ============================================================== bidirectional path tracing
// BidirPathRenderer::render (bidirpath.h:316-395) for one lane per pixel and
// its samples in sample order: a light subpath (traceLightSubpath,
// bidirpath.h:196-219) whose vertices go to a per-lane record array in HBM,
// then a camera subpath (traceCameraSubpath, bidirpath.h:221-257) whose every
// surface vertex connects to every stored light vertex (bdptConnect,
// bidirpath.h:79-123); t = 1 connections (light vertex -> camera vertex) are
// FP64 atomic splats, the rest sum into the pixel's FP64 base film.
//
// Balance-heuristic weights (computeMisWeight, bidirpath.h:125-194, power 1):
// 1 / w = 1 + sum of the light-side ratio products + sum of the camera-side
// ones.  With b_i = (camera-side pdf / light-side pdf) of light vertex i and
// a_j = (light-side pdf / camera-side pdf) of camera vertex j, the light sum of
// strategy (s, t) is b_s'(1 + b_{s-1}'(1 + E_{s-2})) with E_k = b_k (1 + E_{k-1}),
// E_0 = 0, and the camera sum a_t'(1 + a_{t-1}'(1 + D_{t-2})) with
// D_k = a_k (1 + D_{k-1}), D_1 = 0 (the camera vertex never enters); primes
// are the pdfs re-evaluated at the connection, the rest are fixed once the
// subpath has moved two vertices on.  Each connection therefore costs O(1)
// instead of the reference's loop over the path.
//
// Light vertex record (4 float4 per vertex, one plane per field, lane-major):
//   0: position, pdf with which the light subpath made it (area)
//   1: geometric normal, -
//   2: throughput (mPathContrib), E_{k-2}
//   3: direction to the previous light vertex, material (bits)
constexpr int kBdptMaxVertices = 64;
constexpr uint32_t kBdptKeySalt = 0x3c6ef372u;  // the light subpath's walk draws from its own stream

__device__ __forceinline__ float bd_conv(float3 from, float3 to, float3 to_n) {
    const float3 u = to - from;
    const float l2 = dot3(u, u);
    return fabsf(dot3(to_n, u)) / (sqrtf(l2) * l2);
}
__device__ __forceinline__ float bd_cam_pdfW(const DCamera &cam, float3 w) {
    const float c = cam_to_local(cam, w).z;
    return cam.per_pixel_norm / (c * c * c);
}
__device__ __forceinline__ float bd_weight(float inv) {
    const float w = 1.f / inv;
    return w >= 0.f && w <= 1.f ? w : 0.f;  // a zero forward pdf (grazing FP32 hit) weighs 0
}

// Capped at 128 VGPRs (4 waves per SIMD; 328 B/lane spill to scratch): 227 VGPRs
// and 2 waves measured 676 / 4,351 Mrays/s on the 1.31M-triangle frame at 4 spp /
// a 512^2 Cornell box at 16 spp, 4 waves 912-945 / 5,341, 6 waves 1,046 / 3,530,
// 8 waves 803 / 2,932; the film is bit-identical across all of them.
__global__ __launch_bounds__(kBlock, 4) void k_bdpt(DScene s, uint32_t wh, uint32_t spp, uint32_t first_sample,
                                                 uint32_t seed, int N, double *film_base, double *film_splat,
                                                 float4 *__restrict__ recs, unsigned long long *counts, uint2 *spill) {
    __shared__ uint64_t st_ent[kLdsStackDepth * kBlock];
    LaneStack<kBlock> st(st_ent + lane_id(), reinterpret_cast<uint32_t *>(st_ent) + lane_id(), spill);
    TraversalCounters cnt;
    const DCamera &cam = s.cam;
    const float3 cz = f3(cam.zx, cam.zy, cam.zz);
    const size_t lanes = size_t(gridDim.x) * kBlock, lane = size_t(blockIdx.x) * kBlock + lane_id();
    auto rec = [&](int k, int f) -> float4 & { return recs[(size_t(k - 1) * 4 + f) * lanes + lane]; };
    unsigned long long n_ext = 0, n_shd = 0;
    auto occluded = [&](float3 o, float3 d, float len) {
        n_shd++;
        float t = len - kRayEps;
        return traverse_rounds<true, kBlock>(Geo4{s.fnodes, s.tris}, setup_ray(o, d, kRayEps), t, st, cnt) != kInvalid;
    };
    auto closest = [&](float3 o, float3 d, float &t) {
        n_ext++;
        t = INFINITY;
        return traverse_rounds<false, kBlock>(Geo4{s.fnodes, s.tris}, setup_ray(o, d, kRayEps), t, st, cnt);
    };
    for (uint32_t pixel = blockIdx.x * kBlock + lane_id(); pixel < wh; pixel += gridDim.x * kBlock) {
        const uint32_t px = pixel % uint32_t(cam.W), py = pixel / uint32_t(cam.W);
        double acc_r = 0.0, acc_g = 0.0, acc_b = 0.0;
        for (uint32_t k = 0; k < spp; k++) {
            const uint32_t key = path_key(seed, pixel, first_sample + k);
            const uint32_t kl = mix32(key ^ kBdptKeySalt);
            const float3 lens = lens_point(cam, key);
            const float3 cpos = cam_to_world(cam, lens) + f3(cam.ox, cam.oy, cam.oz);
            // ---- light subpath
            const LightSample l = sample_light<false>(s, u01(key, kDimLight), u01(key, kDimLight + 1), key);
            rec(1, 0) = make_float4(l.pos.x, l.pos.y, l.pos.z, l.pdfA);
            rec(1, 1) = make_float4(l.n.x, l.n.y, l.n.z, 0.f);
            rec(1, 2) = make_float4(l.contrib.x, l.contrib.y, l.contrib.z, 0.f);
            int nL = 1;
            // t = 1 connection of vertex sN (light or surface) to the camera vertex
            auto splat_camera = [&](float3 p, float3 n, float3 T, float pf, bool is_light, const DMat &m, float3 wi_w,
                                    float3 q_p, float3 q_n, float q_pf, float E2, int sN) {
                const float3 uv = cpos - p;
                const float len2 = dot3(uv, uv), len = sqrtf(len2);
                const float3 nuv = uv * (1.f / len);
                const float G = fabsf(dot3(n, nuv)) * fabsf(dot3(cz, nuv)) / len2;
                if (G == 0.f) return;
                float rx = 0.f, ry = 0.f;
                const float we1 = camera_we1(cam, lens, -nuv, rx, ry);
                if (we1 == 0.f) return;
                float3 fl;
                const Frame fr = make_frame(n);
                if (is_light) {
                    if (to_local(fr, nuv).y < 0.f) return;
                    fl = f3(kInvPi, kInvPi, kInvPi);
                } else {
                    fl = bsdf_eval(m, to_local(fr, wi_w), to_local(fr, nuv));
                    if (is_black(fl)) return;
                }
                if (occluded(p, nuv, len)) return;
                // weight: light-side ratios only (t = 1)
                const float bs = bd_cam_pdfW(cam, -nuv) * fabsf(dot3(n, nuv)) / len2 / pf;
                float inv;
                if (sN == 1) {
                    inv = 1.f + bs;
                } else {
                    const float cq = bsdf_pdf(m, to_local(fr, nuv), to_local(fr, wi_w)) * bd_conv(p, q_p, q_n);
                    inv = 1.f + bs * (1.f + (cq / q_pf) * (1.f + E2));
                }
                const float c = bd_weight(inv) * G * we1;
                const float3 v = T * fl * c;
                double *dst = film_splat + 3 * (size_t(uint32_t(ry)) * uint32_t(cam.W) + uint32_t(rx));
                atomicAdd(dst + 0, double(v.x));
                atomicAdd(dst + 1, double(v.y));
                atomicAdd(dst + 2, double(v.z));
            };
            const DMat none{};
            splat_camera(l.pos, l.n, l.contrib, l.pdfA, true, none, f3(0, 0, 0), f3(0, 0, 0), f3(0, 0, 0), 1.f, 0.f, 1);
            if (N - 1 >= 2) {
                const Frame lf = make_frame(l.n);
                float3 ro = l.pos;
                float3 rd = to_world(lf, cosine_hemisphere(u01(kl, kDimPixel), u01(kl, kDimPixel + 1)));
                float pdfW = fmaxf(to_local(lf, rd).y, 0.f) * kInvPi;
                float3 w = f3(1.f, 1.f, 1.f), T = l.contrib;
                // the previous vertex q: position, normal, forward pdf
                float3 q_p = l.pos, q_n = l.n;
                float q_pf = l.pdfA;
                float E2 = 0.f;  // E_{k-2} for the vertex about to be made
                for (uint32_t bounce = 0;; bounce++) {
                    float t;
                    const uint32_t tri = closest(ro, rd, t);
                    if (tri == kInvalid) break;
                    const float4 v0 = s.tris[3 * size_t(tri) + 0], v1 = s.tris[3 * size_t(tri) + 1],
                                 v2 = s.tris[3 * size_t(tri) + 2];
                    const int2 gi = s.geom_info[__float_as_uint(v1.w)];
                    if (gi.y >= 0) break;  // a light ends the light subpath
                    const float3 p = ro + rd * t;
                    const float3 ng = normalize3(cross3(xyz(v1) - xyz(v0), xyz(v2) - xyz(v0)));
                    const float3 wiw = -rd;
                    T = T * w;
                    const float pf = pdfW * bd_conv(q_p, p, ng);
                    nL++;
                    rec(nL, 0) = make_float4(p.x, p.y, p.z, pf);
                    rec(nL, 1) = make_float4(ng.x, ng.y, ng.z, 0.f);
                    rec(nL, 2) = make_float4(T.x, T.y, T.z, E2);
                    rec(nL, 3) = make_float4(wiw.x, wiw.y, wiw.z, __int_as_float(gi.x));
                    const DMat m = s.mats[gi.x];
                    splat_camera(p, ng, T, pf, false, m, wiw, q_p, q_n, q_pf, E2, nL);
                    if (nL == N - 1) break;
                    const Frame f = make_frame(ng);
                    const float3 wi = to_local(f, wiw);
                    const uint32_t dim = kDimBounce0 + kDimsPerBounce * bounce + kBdBsdf;
                    float3 wo;
                    float pdf = 0.f;
                    w = bsdf_sample(m, wi, u01(kl, dim), u01(kl, dim + 1), wo, pdf);
                    if (is_black(w)) break;
                    // the previous vertex's camera-side pdf given this direction -> E_{k-1}
                    const float pr_q = bsdf_pdf(m, wo, wi) * bd_conv(p, q_p, q_n);
                    E2 = (pr_q / q_pf) * (1.f + E2);  // E_{k-1}, stored with the next vertex
                    q_p = p; q_n = ng; q_pf = pf;
                    pdfW = pdf;
                    ro = p;
                    rd = to_world(f, wo);
                }
            }
            // ---- camera subpath
            float pdfW;
            float3 rd = primary_dir(cam, key, px, py, lens, pdfW);
            float3 ro = cpos;
            float3 q_p = cpos, q_n = cz;  // previous camera-subpath vertex
            float q_pf = 1.f;
            float3 T = f3(1.f, 1.f, 1.f), w = f3(1.f, 1.f, 1.f);
            float D2 = 0.f;  // D_{t-2}
            float sr = 0.f, sg = 0.f, sb = 0.f;
            for (int tv = 2; tv <= N; tv++) {
                float th;
                const uint32_t tri = closest(ro, rd, th);
                if (tri == kInvalid) break;
                const float4 v0 = s.tris[3 * size_t(tri) + 0], v1 = s.tris[3 * size_t(tri) + 1],
                             v2 = s.tris[3 * size_t(tri) + 2];
                const int2 gi = s.geom_info[__float_as_uint(v1.w)];
                const float3 p = ro + rd * th;
                const float3 ng = normalize3(cross3(xyz(v1) - xyz(v0), xyz(v2) - xyz(v0)));
                const float3 wiw = -rd;
                T = T * w;
                const float pf = pdfW * bd_conv(q_p, p, ng);
                if (gi.y >= 0) {
                    // s = 0: the camera subpath ends on an emitter (Le1 * Le0 = L on its front)
                    const float cl = dot3(ng, wiw);
                    if (cl >= 0.f) {
                        const DLight &L = s.lights[gi.y];
                        const float at = L.pdf_select * L.inv_area / pf;
                        float inv;
                        if (tv == 2) {
                            inv = 1.f + at;
                        } else {
                            const float pr_q = cl * kInvPi * bd_conv(p, q_p, q_n);
                            inv = 1.f + at * (1.f + (pr_q / q_pf) * (1.f + D2));
                        }
                        const float c = bd_weight(inv) * (kInvPi * kPi);
                        sr += T.x * L.L[0] * c; sg += T.y * L.L[1] * c; sb += T.z * L.L[2] * c;
                    }
                    break;
                }
                if (tv == N) break;
                const DMat m = s.mats[gi.x];
                const Frame f = make_frame(ng);
                const float3 wi = to_local(f, wiw);
                // connections to the stored light vertices, s + t <= N
                const int smax = min(nL, N - tv);
                for (int sN = 1; sN <= smax; sN++) {
                    const float4 A = rec(sN, 0), B = rec(sN, 1), Cc = rec(sN, 2);
                    const float3 lp = xyz(A), ln = xyz(B);
                    const float3 uv = p - lp;  // light vertex -> camera vertex
                    const float len2 = dot3(uv, uv), len = sqrtf(len2);
                    const float3 nuv = uv * (1.f / len);
                    const float G = fabsf(dot3(ln, nuv)) * fabsf(dot3(ng, nuv)) / len2;
                    if (G == 0.f) continue;
                    const float3 wo_c = to_local(f, -nuv);
                    const float3 fc = bsdf_eval(m, wi, wo_c);
                    if (is_black(fc)) continue;
                    const Frame lfr = make_frame(ln);
                    float3 fl;
                    float lpdf_c;  // light-side pdf of the camera vertex (solid angle at the light vertex)
                    DMat lm{};
                    float3 lwi_l = f3(0, 0, 0);
                    if (sN == 1) {
                        const float cl = to_local(lfr, nuv).y;
                        if (cl < 0.f) continue;
                        fl = f3(kInvPi, kInvPi, kInvPi);
                        lpdf_c = cl * kInvPi;
                    } else {
                        const float4 Dd = rec(sN, 3);
                        lm = s.mats[__float_as_int(Dd.w)];
                        lwi_l = to_local(lfr, xyz(Dd));
                        const float3 lwo = to_local(lfr, nuv);
                        fl = bsdf_eval(lm, lwi_l, lwo);
                        if (is_black(fl)) continue;
                        lpdf_c = bsdf_pdf(lm, lwi_l, lwo);
                    }
                    if (occluded(lp, nuv, len)) continue;
                    // light-side sum
                    const float bs = bsdf_pdf(m, wi, wo_c) * fabsf(dot3(ln, nuv)) / len2 / A.w;
                    float lsum;
                    if (sN == 1) {
                        lsum = bs;
                    } else {
                        const float4 Aq = rec(sN - 1, 0), Bq = rec(sN - 1, 1);
                        const float cq = bsdf_pdf(lm, to_local(lfr, nuv), lwi_l) * bd_conv(lp, xyz(Aq), xyz(Bq));
                        lsum = bs * (1.f + (cq / Aq.w) * (1.f + Cc.w));
                    }
                    // camera-side sum
                    const float at = lpdf_c * fabsf(dot3(ng, nuv)) / len2 / pf;
                    float csum;
                    if (tv == 2) {
                        csum = at;
                    } else {
                        const float lq = bsdf_pdf(m, wo_c, wi) * bd_conv(p, q_p, q_n);
                        csum = at * (1.f + (lq / q_pf) * (1.f + D2));
                    }
                    const float c = bd_weight(1.f + lsum + csum) * G;
                    const float3 v = xyz(Cc) * fl * fc * T * c;
                    sr += v.x; sg += v.y; sb += v.z;
                }
                // continue the camera subpath
                const uint32_t dim = kDimBounce0 + kDimsPerBounce * uint32_t(tv - 2) + kBdBsdf;
                float3 wo;
                float pdf = 0.f;
                w = bsdf_sample(m, wi, u01(key, dim), u01(key, dim + 1), wo, pdf);
                if (is_black(w)) break;
                if (tv >= 3) {  // D_{t-1} from the previous vertex's light-side pdf given this direction
                    const float pr_q = bsdf_pdf(m, wo, wi) * bd_conv(p, q_p, q_n);
                    D2 = (pr_q / q_pf) * (1.f + D2);
                }
                q_p = p; q_n = ng; q_pf = pf;
                pdfW = pdf;
                ro = p;
                rd = to_world(f, wo);
            }
            acc_r += sr; acc_g += sg; acc_b += sb;
        }
        double *dst = film_base + 3 * size_t(pixel);
        dst[0] += acc_r; dst[1] += acc_g; dst[2] += acc_b;
    }
    atomicAdd(counts + 0, n_ext);
    atomicAdd(counts + 1, n_shd);
}

// Resident blocks per CU of a kernel (HIP occupancy API), cached per device.
static int occupancy(wr_device *d, const void *fn, int block) {
    auto it = d->occ_cache.find(fn);
    if (it != d->occ_cache.end()) return it->second;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, block, 0) != hipSuccess || nb <= 0) nb = 1;
    nb = std::min(nb, kMaxTraversalBlocksPerCU);
    d->occ_cache[fn] = nb;
    return nb;
}

// The entry points below inherit C linkage from their declarations in wurst_amd.h.

const char *wr_last_error(void) { return wr::error_string().c_str(); }
const char *wr_version(void) { return "wurst_amd 0.1 (gfx950, BVH4-128B, wavefront NEE+MIS)"; }

static int scratch_setup(wr_device *d);

int wr_device_create(int ordinal, wr_device **out) {
    if (!out) return set_error(WR_ERR_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    int count = 0;
    WR_HIP(hipGetDeviceCount(&count));
    if (ordinal < 0 || ordinal >= count) return set_error(WR_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    WR_HIP(hipSetDevice(ordinal));
    hipDeviceProp_t prop;
    WR_HIP(hipGetDeviceProperties(&prop, ordinal));
    wr_device *d = new wr_device;
    d->ordinal = ordinal;
    d->n_cu = prop.multiProcessorCount;
    auto occ = [](const void *fn, int block, int &out_blocks) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, block, 0) == hipSuccess && nb > 0) out_blocks = nb;
        out_blocks = std::min(out_blocks, kMaxTraversalBlocksPerCU);
    };
    occ(reinterpret_cast<const void *>(&k_trace_bounce_fast), kBlock, d->occ_rays);
    occ(reinterpret_cast<const void *>(&k_shade<false>), kShadeBlock, d->occ_shade);

    occ(reinterpret_cast<const void *>(&k_regen<false>), kShadeBlock, d->occ_generate);
    occ(reinterpret_cast<const void *>(&k_trace<false>), kBlock, d->occ_trace);
    if (int rc = scratch_setup(d)) {
        delete d;
        return rc;
    }
    *out = d;
    return WR_OK;
}

void wr_device_destroy(wr_device *dev) { delete dev; }

void wr_scene_destroy(wr_scene *s) {
    if (!s) return;
    (void)hipSetDevice(s->dev->ordinal);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    (void)hipFree(s->d_nodes);
    (void)hipFree(s->d_qnodes);
    (void)hipFree(s->d_tris);
    (void)hipFree(s->d_ltris);
    (void)hipFree(s->d_geom);
    (void)hipFree(s->d_mats);
    (void)hipFree(s->d_lights);
    (void)hipFree(s->d_env);
    (void)hipFree(s->d_env_cdf);
    (void)hipFree(s->d_spill);
    (void)hipFree(s->d_bdpt);
    (void)hipFree(s->pool);
    if (s->h_counters) (void)hipHostFree(s->h_counters);
    for (auto &e : s->ev)
        if (e) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

static int validate_desc(const wr_scene_desc *d) {
    if (!d) return set_error(WR_ERR_INVALID_ARGUMENT, "desc is null");
    if (d->n_tris > 0 && (!d->verts || !d->tris || !d->tri_geom))
        return set_error(WR_ERR_INVALID_ARGUMENT, "null geometry arrays");
    if (d->n_tris >= kMaxTriangles) return set_error(WR_ERR_UNSUPPORTED, "too many triangles (limit 2^28)");
    if (d->n_verts > 0 && !d->verts) return set_error(WR_ERR_INVALID_ARGUMENT, "null vertex array");
    for (size_t i = 0; i < 3 * size_t(d->n_verts); i++)
        if (!std::isfinite(d->verts[i])) return set_error(WR_ERR_INVALID_ARGUMENT, "non-finite vertex coordinate");
    if (d->camera.width <= 0 || d->camera.height <= 0) return set_error(WR_ERR_INVALID_ARGUMENT, "bad resolution");
    if ((uint64_t)d->camera.width * d->camera.height >= (1ull << 31))
        return set_error(WR_ERR_UNSUPPORTED, "resolution too large");
    if (!(d->camera.focal_distance > 0.f))
        return set_error(WR_ERR_INVALID_ARGUMENT, "focal distance must be > 0 (the JSON key \"radius\" sets it)");
    if (!(d->camera.fovy_deg > 0.f && d->camera.fovy_deg < 180.f))
        return set_error(WR_ERR_INVALID_ARGUMENT, "fovy must be in (0, 180) degrees");
    for (uint32_t i = 0; i < 3 * d->n_tris; i++)
        if (d->tris[i] >= d->n_verts) return set_error(WR_ERR_INVALID_ARGUMENT, "triangle vertex index out of range");
    uint32_t prev = 0;
    for (uint32_t t = 0; t < d->n_tris; t++) {
        if (d->tri_geom[t] >= d->n_geoms) return set_error(WR_ERR_INVALID_ARGUMENT, "tri_geom out of range");
        if (d->tri_geom[t] < prev) return set_error(WR_ERR_INVALID_ARGUMENT, "triangles of a geometry must be contiguous and ordered");
        prev = d->tri_geom[t];
    }
    for (uint32_t g = 0; g < d->n_geoms; g++) {
        const int m = d->geom_material[g], l = d->geom_light[g];
        if (l >= (int)d->n_lights) return set_error(WR_ERR_INVALID_ARGUMENT, "geom_light out of range");
        if (l < 0 && (m < 0 || m >= (int)d->n_materials))
            return set_error(WR_ERR_INVALID_ARGUMENT, "geometry without a valid material");
    }
    for (uint32_t m = 0; m < d->n_materials; m++)
        if (d->materials[m].type < 0 || d->materials[m].type > 3) return set_error(WR_ERR_INVALID_ARGUMENT, "unknown material type");
    for (uint32_t l = 0; l < d->n_lights; l++)
        if (d->lights[l].geom < 0 || (uint32_t)d->lights[l].geom >= d->n_geoms)
            return set_error(WR_ERR_INVALID_ARGUMENT, "light geometry out of range");
    return WR_OK;
}

int wr_scene_validate(const wr_scene_desc *desc) { return validate_desc(desc); }

int wr_bvh_selftest(const wr_scene_desc *d, uint64_t *n_nodes, uint32_t *max_stack, uint32_t *depth) {
    int rc = validate_desc(d);
    if (rc) return rc;
    BvhBuildResult bvh;
    if (!build_bvh4(d->verts, d->n_verts, d->tris, d->n_tris, d->tri_geom, &bvh))
        return set_error(WR_ERR_OUT_OF_MEMORY, "host allocation failed during BVH build");
    const std::string err = verify_bvh4(bvh, d->verts, d->tris, d->n_tris);
    if (n_nodes) *n_nodes = bvh.n_nodes;
    if (max_stack) *max_stack = bvh.max_stack;
    if (depth) *depth = bvh.depth;
    free_bvh(&bvh);
    if (!err.empty()) return set_error(WR_ERR_DEVICE, "BVH invariant violated: " + err);
    return WR_OK;
}

template <typename T>
static int upload(T **dst, const void *src, size_t bytes, uint64_t &total) {
    WR_HIP(hipMalloc(reinterpret_cast<void **>(dst), std::max<size_t>(bytes, 128)));
    if (bytes) WR_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    total += bytes;
    return WR_OK;
}

// Host SAH build, uploaded.
static int bvh_host(wr_scene *s, const wr_scene_desc *d) {
    const auto t0 = std::chrono::steady_clock::now();
    BvhBuildResult bvh;
    if (!build_bvh4(d->verts, d->n_verts, d->tris, d->n_tris, d->tri_geom, &bvh))
        return set_error(WR_ERR_OUT_OF_MEMORY, "host allocation failed during BVH build");
    s->leaf_to_orig.resize(d->n_tris);
    for (uint32_t i = 0; i < d->n_tris; i++) s->leaf_to_orig[i] = bvh.tris[i].orig;
    s->n_nodes = bvh.n_nodes;
    s->max_stack = bvh.max_stack;
    s->bvh_depth = bvh.depth;
    int rc = upload(&s->d_nodes, bvh.nodes, sizeof(Node4) * bvh.n_nodes, s->device_bytes);
    if (!rc) rc = upload(&s->d_tris, bvh.tris, sizeof(TriRecord) * bvh.n_tris, s->device_bytes);
    free_bvh(&bvh);
    s->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

// GPU PLOC build from a temporary device copy of the mesh.
static int bvh_gpu(wr_scene *s, const wr_scene_desc *d) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t n = d->n_tris;
    float *dv = nullptr;
    uint32_t *dt = nullptr, *dg = nullptr;
    uint64_t mesh_bytes = 0;
    GpuBvhResult g;
    int rc = upload(&dv, d->verts, sizeof(float) * 3 * size_t(d->n_verts), mesh_bytes);
    if (!rc) rc = upload(&dt, d->tris, sizeof(uint32_t) * 3 * size_t(n), mesh_bytes);
    if (!rc) rc = upload(&dg, d->tri_geom, sizeof(uint32_t) * size_t(n), mesh_bytes);
    s->leaf_to_orig.resize(n);
    if (!rc) rc = build_bvh4_gpu(dv, dt, dg, n, s->stream, s->ploc_radius, s->leaf_to_orig.data(), &g);
    (void)hipFree(dv);
    (void)hipFree(dt);
    (void)hipFree(dg);
    s->d_nodes = g.nodes;
    s->d_tris = reinterpret_cast<float4 *>(g.tris);
    if (rc) return rc;
    s->device_bytes += sizeof(Node4) * uint64_t(g.n_nodes) + sizeof(TriRecord) * uint64_t(n);
    s->n_nodes = g.n_nodes;
    s->max_stack = g.max_stack;
    s->bvh_depth = g.depth;
    s->bvh_iterations = g.iterations;
    s->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return WR_OK;
}

static int scene_create_impl(wr_device *dev, const wr_scene_desc *d, int builder, int radius, wr_scene *s) {
    s->dev = dev;
    int rc = set_device(dev);
    if (rc) return rc;
    WR_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    WR_HIP(hipEventCreate(&s->ev[0]));
    WR_HIP(hipEventCreate(&s->ev[1]));
    s->W = d->camera.width;
    s->H = d->camera.height;
    for (int k = 0; k < 3; k++) { s->box[k] = INFINITY; s->box[3 + k] = -INFINITY; }
    for (uint32_t v = 0; v < d->n_verts; v++)
        for (int k = 0; k < 3; k++) {
            s->box[k] = std::min(s->box[k], d->verts[3 * size_t(v) + k]);
            s->box[3 + k] = std::max(s->box[3 + k], d->verts[3 * size_t(v) + k]);
        }
    // host tables
    s->tri_geom.assign(d->tri_geom, d->tri_geom + d->n_tris);
    s->geom_first.assign(d->n_geoms, kInvalid);
    for (uint32_t t = 0; t < d->n_tris; t++)
        if (s->geom_first[d->tri_geom[t]] == kInvalid) s->geom_first[d->tri_geom[t]] = t;
    // BVH
    s->builder = builder;
    s->ploc_radius = radius;
    s->n_tris = d->n_tris;
    rc = builder == WR_BVH_HOST ? bvh_host(s, d) : bvh_gpu(s, d);
    if (rc) return rc;
    WR_HIP(hipMalloc(&s->d_qnodes, sizeof(NodeQ) * std::max<uint64_t>(s->n_nodes, 1)));
    s->device_bytes += sizeof(NodeQ) * s->n_nodes;
    if (s->n_nodes) {
        k_quantize_nodes<<<uint32_t((s->n_nodes + 255) / 256), 256, 0, s->stream>>>(s->d_nodes, s->d_qnodes,
                                                                                    uint32_t(s->n_nodes));
        WR_HIP(hipGetLastError());
        WR_HIP(hipStreamSynchronize(s->stream));
    }
    if (s->max_stack > (uint32_t)kStackDepth) return set_error(WR_ERR_UNSUPPORTED, "BVH deeper than the traversal stack");
    s->max_trav_blocks = dev->n_cu * kMaxTraversalBlocksPerCU;
    {
        const size_t spill_bytes = sizeof(uint2) * size_t(s->max_trav_blocks) * kBlock * kSpillDepth;
        WR_HIP(hipMalloc(&s->d_spill, spill_bytes));
        s->device_bytes += spill_bytes;
    }
    // geometry info
    std::vector<int2> gi(d->n_geoms);
    for (uint32_t g = 0; g < d->n_geoms; g++) gi[g] = make_int2(d->geom_material[g], d->geom_light[g]);
    rc = upload(&s->d_geom, gi.data(), sizeof(int2) * gi.size(), s->device_bytes);
    if (rc) return rc;
    std::vector<DMat> mats(d->n_materials);
    for (uint32_t m = 0; m < d->n_materials; m++) {
        const wr_material &src = d->materials[m];
        DMat &dm = mats[m];
        dm.type = src.type;
        dm.twosided = src.twosided;
        dm.sigma = src.sigma;
        dm.n = src.exponent;
        for (int k = 0; k < 3; k++) { dm.rho_d[k] = src.rho_d[k]; dm.rho_s[k] = src.rho_s[k]; }
        dm.rho_d[3] = dm.rho_s[3] = 0.f;
    }
    rc = upload(&s->d_mats, mats.data(), sizeof(DMat) * mats.size(), s->device_bytes);
    if (rc) return rc;
    // lights: selection CdfTable over |L|_2, per-light triangle-area CdfTable (FP64 on the host)
    std::vector<DLight> lights(d->n_lights);
    std::vector<float4> ltris;
    std::vector<double> w(d->n_lights);
    double wsum = 0.0;
    for (uint32_t l = 0; l < d->n_lights; l++) {
        const float *L = d->lights[l].L;
        w[l] = std::sqrt(double(L[0]) * L[0] + double(L[1]) * L[1] + double(L[2]) * L[2]);
        wsum += w[l];
    }
    if (d->n_lights > 0 && !(wsum > 0.0)) return set_error(WR_ERR_INVALID_ARGUMENT, "all lights have zero radiance");
    s->light_w = w;
    double cdf = 0.0;
    for (uint32_t l = 0; l < d->n_lights; l++) {
        DLight &dl = lights[l];
        const uint32_t g = (uint32_t)d->lights[l].geom;
        const double pdf = w[l] / wsum;
        cdf += pdf;
        dl.pdf_select = float(pdf);
        dl.cdf_select = (l + 1 == d->n_lights) ? 1.0f : float(cdf);
        for (int k = 0; k < 3; k++) dl.L[k] = d->lights[l].L[k];
        uint32_t first = s->geom_first[g];
        if (first == kInvalid) return set_error(WR_ERR_INVALID_ARGUMENT, "light geometry has no triangles");
        uint32_t cnt = 0;
        while (first + cnt < d->n_tris && d->tri_geom[first + cnt] == g) cnt++;
        std::vector<double> area(cnt);
        double asum = 0.0;
        for (uint32_t i = 0; i < cnt; i++) {
            const uint32_t *t = d->tris + 3 * size_t(first + i);
            const float *a = d->verts + 3 * size_t(t[0]), *b = d->verts + 3 * size_t(t[1]), *c = d->verts + 3 * size_t(t[2]);
            double e1[3], e2[3];
            for (int k = 0; k < 3; k++) { e1[k] = double(b[k]) - a[k]; e2[k] = double(c[k]) - a[k]; }
            const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
            area[i] = std::sqrt(cx * cx + cy * cy + cz * cz) / 2.0;
            asum += area[i];
        }
        if (!(asum > 0.0)) return set_error(WR_ERR_INVALID_ARGUMENT, "light with zero area");
        dl.area = float(asum);
        dl.inv_area = float(1.0 / asum);
        dl.first_ltri = uint32_t(ltris.size() / 3);
        dl.n_ltri = cnt;
        double tc = 0.0;
        for (uint32_t i = 0; i < cnt; i++) {
            const uint32_t *t = d->tris + 3 * size_t(first + i);
            const double tp = area[i] / asum;
            tc += tp;
            const float tcdf = (i + 1 == cnt) ? 1.0f : float(tc);
            const float *a = d->verts + 3 * size_t(t[0]), *b = d->verts + 3 * size_t(t[1]), *c = d->verts + 3 * size_t(t[2]);
            ltris.push_back(make_float4(a[0], a[1], a[2], tcdf));
            ltris.push_back(make_float4(b[0], b[1], b[2], float(tp)));
            ltris.push_back(make_float4(c[0], c[1], c[2], 0.f));
        }
    }
    s->h_lights = lights;
    rc = upload(&s->d_lights, lights.data(), sizeof(DLight) * lights.size(), s->device_bytes);  // >= 128 B allocated
    if (!rc) rc = upload(&s->d_ltris, ltris.data(), sizeof(float4) * ltris.size(), s->device_bytes);
    if (rc) return rc;
    // camera (ThinlensCamera ctor, thinlens.h:40-55; Camera ctor, camera.h:9-17), FP64 then rounded
    const wr_camera &c = d->camera;
    double o[3] = {c.pos[0], c.pos[1], c.pos[2]}, la[3] = {c.lookat[0], c.lookat[1], c.lookat[2]},
           up[3] = {c.up[0], c.up[1], c.up[2]};
    auto norm = [](double v[3]) {
        double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int k = 0; k < 3; k++) v[k] /= l;
    };
    auto cross = [](const double a[3], const double b[3], double r[3]) {
        r[0] = a[1] * b[2] - a[2] * b[1];
        r[1] = a[2] * b[0] - a[0] * b[2];
        r[2] = a[0] * b[1] - a[1] * b[0];
    };
    double z[3] = {la[0] - o[0], la[1] - o[1], la[2] - o[2]}, x[3], y[3];
    norm(z);
    cross(up, z, x);
    norm(x);
    cross(z, x, y);
    norm(y);
    DCamera &dc = s->ds.cam;
    dc.ox = c.pos[0]; dc.oy = c.pos[1]; dc.oz = c.pos[2];
    dc.xx = float(x[0]); dc.xy = float(x[1]); dc.xz = float(x[2]);
    dc.yx = float(y[0]); dc.yy = float(y[1]); dc.yz = float(y[2]);
    dc.zx = float(z[0]); dc.zy = float(z[1]); dc.zz = float(z[2]);
    const double fovy = double(c.fovy_deg) * 3.14159265358979323846 / 180.0;
    const double dist = double(c.height) * 0.5 / std::tan(fovy * 0.5);
    dc.d = float(dist);
    dc.W = c.width;
    dc.H = c.height;
    dc.invW = float(1.0 / c.width);
    dc.invH = float(1.0 / c.height);
    dc.per_pixel_norm = float(dist * dist / (double(c.width) * double(c.height)));
    dc.per_film_norm = float(dist * dist);
    dc.aperture = c.aperture_radius;
    dc.focal = c.focal_distance;
    s->ds.nodes = s->d_qnodes;
    s->ds.fnodes = s->d_nodes;
    s->ds.tris = s->d_tris;
    s->ds.geom_info = s->d_geom;
    s->ds.mats = s->d_mats;
    s->ds.lights = s->d_lights;
    s->ds.ltris = s->d_ltris;
    s->ds.n_lights = d->n_lights;
    return WR_OK;
}

int wr_scene_create(wr_device *dev, const wr_scene_desc *desc, wr_scene **out) {
    return wr_scene_create_ex(dev, desc, nullptr, out);
}

int wr_scene_create_ex(wr_device *dev, const wr_scene_desc *desc, const wr_build_params *params, wr_scene **out) {
    if (!dev || !out) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const int builder = params ? params->builder : WR_BVH_GPU;
    const int radius = params ? params->ploc_radius : 0;
    if (builder != WR_BVH_GPU && builder != WR_BVH_HOST) return set_error(WR_ERR_INVALID_ARGUMENT, "unknown BVH builder");
    if (radius < 0 || radius > 64) return set_error(WR_ERR_INVALID_ARGUMENT, "ploc_radius must be in [0, 64]");
    int rc = validate_desc(desc);
    if (rc) return rc;
    wr_scene *s = new wr_scene;
    rc = scene_create_impl(dev, desc, builder, radius, s);
    if (rc) {
        std::string msg = wr::error_string();
        wr_scene_destroy(s);
        wr::error_string() = msg;
        return rc;
    }
    *out = s;
    return WR_OK;
}

int wr_scene_info(const wr_scene *s, uint64_t *n_nodes, uint64_t *n_tris, uint32_t *max_stack, uint64_t *bytes) {
    if (!s) return set_error(WR_ERR_INVALID_ARGUMENT, "scene is null");
    if (n_nodes) *n_nodes = s->n_nodes;
    if (n_tris) *n_tris = s->n_tris;
    if (max_stack) *max_stack = s->max_stack;
    if (bytes) *bytes = s->device_bytes;
    return WR_OK;
}

int wr_scene_bvh_check(wr_scene *s, const wr_scene_desc *d, wr_bvh_stats *st) {
    if (!s || !d || !st) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    if (d->n_tris != s->n_tris) return set_error(WR_ERR_INVALID_ARGUMENT, "desc is not the scene's description");
    int rc = set_device(s->dev);
    if (rc) return rc;
    BvhBuildResult r;
    r.n_nodes = uint32_t(s->n_nodes);
    r.n_tris = uint32_t(s->n_tris);
    r.max_stack = s->max_stack;
    r.depth = s->bvh_depth;
    std::vector<Node4> nodes(s->n_nodes);
    std::vector<TriRecord> tris(s->n_tris);
    WR_HIP(hipStreamSynchronize(s->stream));
    WR_HIP(hipMemcpy(nodes.data(), s->d_nodes, sizeof(Node4) * nodes.size(), hipMemcpyDeviceToHost));
    if (!tris.empty())
        WR_HIP(hipMemcpy(tris.data(), s->d_tris, sizeof(TriRecord) * tris.size(), hipMemcpyDeviceToHost));
    r.nodes = nodes.data();
    r.tris = tris.data();
    const std::string err = verify_bvh4(r, d->verts, d->tris, d->n_tris);
    if (!err.empty()) return set_error(WR_ERR_DEVICE, "BVH invariant violated: " + err);
    for (uint32_t i = 0; i < s->n_tris; i++)
        if (tris[i].geom != d->tri_geom[tris[i].orig] || tris[i].orig != s->leaf_to_orig[i])
            return set_error(WR_ERR_DEVICE, "BVH invariant violated: leaf triangle geometry id or order map");
    // the traversal copy: same refs, and every decoded child box (the kernels'
    // fmaf(q, scale, origin), IEEE single rounding on both sides) contains the Node4 box
    {
        std::vector<NodeQ> qn(s->n_nodes);
        WR_HIP(hipMemcpy(qn.data(), s->d_qnodes, sizeof(NodeQ) * qn.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < qn.size(); i++) {
            const Node4 &a = nodes[i];
            const NodeQ &b = qn[i];
            const float *lo[3] = {a.lo_x, a.lo_y, a.lo_z}, *hi[3] = {a.hi_x, a.hi_y, a.hi_z};
            const float sc[3] = {b.scale_x, b.scale_y, b.scale_z};
            for (int c = 0; c < 4; c++) {
                if (b.child[c] != a.child[c])
                    return set_error(WR_ERR_DEVICE, "BVH invariant violated: quantised node refs differ");
                if (a.child[c] == kEmptyChild) continue;
                for (int k = 0; k < 3; k++) {
                    const float pl = std::fma(float((b.q[2 * k] >> (8 * c)) & 0xFFu), sc[k], b.org[k]);
                    const float ph = std::fma(float((b.q[2 * k + 1] >> (8 * c)) & 0xFFu), sc[k], b.org[k]);
                    if (!(pl <= lo[k][c] && ph >= hi[k][c]))
                        return set_error(WR_ERR_DEVICE, "BVH invariant violated: quantised box of node " +
                                                            std::to_string(i) + " does not contain its child " +
                                                            std::to_string(c));
                }
            }
        }
    }
    // SAH cost relative to the root box
    auto area = [](const Node4 &n, int c) {
        const double dx = double(n.hi_x[c]) - n.lo_x[c], dy = double(n.hi_y[c]) - n.lo_y[c],
                     dz = double(n.hi_z[c]) - n.lo_z[c];
        return (dx < 0 || dy < 0 || dz < 0) ? 0.0 : dx * dy + dy * dz + dz * dx;
    };
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int c = 0; c < 4; c++)
        if (nodes[0].child[c] != kEmptyChild) {
            lo[0] = std::min(lo[0], double(nodes[0].lo_x[c])); hi[0] = std::max(hi[0], double(nodes[0].hi_x[c]));
            lo[1] = std::min(lo[1], double(nodes[0].lo_y[c])); hi[1] = std::max(hi[1], double(nodes[0].hi_y[c]));
            lo[2] = std::min(lo[2], double(nodes[0].lo_z[c])); hi[2] = std::max(hi[2], double(nodes[0].hi_z[c]));
        }
    const double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    const double aroot = (s->n_tris == 0) ? 0.0 : dx * dy + dy * dz + dz * dx;
    double cost = 1.0;
    if (aroot > 0.0)
        for (const Node4 &n : nodes)
            for (int c = 0; c < 4; c++) {
                if (n.child[c] == kEmptyChild) continue;
                cost += area(n, c) / aroot * ((n.child[c] & kLeafBit) ? double(n.count[c]) : 1.0);
            }
    // FNV-1a over nodes then triangles
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](const void *p, size_t bytes) {
        const uint8_t *b = static_cast<const uint8_t *>(p);
        for (size_t i = 0; i < bytes; i++) h = (h ^ b[i]) * 1099511628211ull;
    };
    mix(nodes.data(), sizeof(Node4) * nodes.size());
    mix(tris.data(), sizeof(TriRecord) * tris.size());
    st->builder = s->builder;
    st->depth = s->bvh_depth;
    st->max_stack = s->max_stack;
    st->iterations = s->bvh_iterations;
    st->n_nodes = s->n_nodes;
    st->build_ms = s->build_ms;
    st->sah_cost = cost;
    st->hash = h;
    return WR_OK;
}

int wr_scene_set_medium(wr_scene *s, const wr_medium *m) {
    if (!s) return set_error(WR_ERR_INVALID_ARGUMENT, "scene is null");
    DMedium dm{};
    if (m) {
        if (!(m->g > -1.f && m->g < 1.f)) return set_error(WR_ERR_INVALID_ARGUMENT, "HG g must be in (-1, 1)");
        for (int c = 0; c < 3; c++) {
            if (!(m->sigma_a[c] >= 0.f) || !(m->sigma_s[c] >= 0.f) || !std::isfinite(m->sigma_a[c] + m->sigma_s[c]))
                return set_error(WR_ERR_INVALID_ARGUMENT, "medium coefficients must be finite and >= 0");
            dm.sig_s[c] = m->sigma_s[c];
            dm.sig_t[c] = m->sigma_a[c] + m->sigma_s[c];
            dm.on |= dm.sig_t[c] > 0.f;
        }
        dm.g = m->g;
    }
    s->ds.med = dm;
    return WR_OK;
}

// The light-selection CdfTable over the area lights' |L|_2 and, last, the
// envmap's weight 1.23 (generalsubpathsampler.h:167-176; envmap.h:22-26).
static int rebuild_light_select(wr_scene *s, bool env) {
    double wsum = env ? 1.23 : 0.0;
    for (double w : s->light_w) wsum += w;
    double cdf = 0.0;
    for (size_t l = 0; l < s->h_lights.size(); l++) {
        const double pdf = s->light_w[l] / wsum;
        cdf += pdf;
        s->h_lights[l].pdf_select = float(pdf);
        s->h_lights[l].cdf_select = (!env && l + 1 == s->h_lights.size()) ? 1.0f : float(cdf);
    }
    s->ds.env_select = env ? float(1.23 / wsum) : 0.f;
    if (!s->h_lights.empty())
        WR_HIP(hipMemcpy(s->d_lights, s->h_lights.data(), sizeof(DLight) * s->h_lights.size(), hipMemcpyHostToDevice));
    return WR_OK;
}

int wr_scene_set_envmap(wr_scene *s, const float *rgb, uint32_t width, uint32_t height) {
    if (!s) return set_error(WR_ERR_INVALID_ARGUMENT, "scene is null");
    int rc = set_device(s->dev);
    if (rc) return rc;
    WR_HIP(hipStreamSynchronize(s->stream));
    if (!rgb || width == 0) {
        (void)hipFree(s->d_env);
        (void)hipFree(s->d_env_cdf);
        s->d_env = nullptr;
        s->d_env_cdf = nullptr;
        s->ds.env_tex = nullptr;
        s->ds.env_cdf = s->ds.env_pdf = nullptr;
        s->ds.env_w = s->ds.env_h = 0;
        return rebuild_light_select(s, false);
    }
    if (height == 0 || width != 2 * height) return set_error(WR_ERR_INVALID_ARGUMENT, "envmap has wrong size");
    if (uint64_t(width) * height > (uint64_t(1) << 26)) return set_error(WR_ERR_UNSUPPORTED, "envmap too large");
    const size_t n = size_t(width) * height;
    std::vector<float4> tex(n);
    std::vector<double> w(n);
    double sum = 0.0;
    for (uint32_t y = 0; y < height; y++) {
        const double sine = std::sin((double(y) + 0.5) / height * 3.14159265358979323846);
        for (uint32_t x = 0; x < width; x++) {
            const size_t i = size_t(y) * width + x;
            const float *t = rgb + 3 * i;
            for (int c = 0; c < 3; c++)
                if (!(t[c] >= 0.f) || !std::isfinite(t[c]))
                    return set_error(WR_ERR_INVALID_ARGUMENT, "envmap texels must be finite and >= 0");
            tex[i] = make_float4(t[0], t[1], t[2], 0.f);
            // EquiRectRep's texel weight: sin(theta) |rgb|_2 (equirectangular.h:40-58)
            w[i] = sine * std::sqrt(double(t[0]) * t[0] + double(t[1]) * t[1] + double(t[2]) * t[2]);
            sum += w[i];
        }
    }
    if (!(sum > 0.0)) return set_error(WR_ERR_INVALID_ARGUMENT, "envmap is black");
    std::vector<float> cp(2 * n);  // cdf, then pdf (CdfTable in FP64, rounded)
    double cdf = 0.0;
    for (size_t i = 0; i < n; i++) {
        const double pdf = w[i] / sum;
        cdf += pdf;
        cp[i] = i + 1 == n ? 1.0f : float(cdf);
        cp[n + i] = float(pdf);
    }
    float4 *d_tex = nullptr;
    float *d_cp = nullptr;
    uint64_t bytes = 0;
    rc = upload(&d_tex, tex.data(), sizeof(float4) * n, bytes);
    if (!rc) rc = upload(&d_cp, cp.data(), sizeof(float) * 2 * n, bytes);
    if (rc) {
        (void)hipFree(d_tex);
        return rc;
    }
    (void)hipFree(s->d_env);
    (void)hipFree(s->d_env_cdf);
    s->d_env = d_tex;
    s->d_env_cdf = d_cp;
    s->ds.env_tex = d_tex;
    s->ds.env_cdf = d_cp;
    s->ds.env_pdf = d_cp + n;
    s->ds.env_w = int(width);
    s->ds.env_h = int(height);
    return rebuild_light_select(s, true);
}

int wr_film_create(wr_scene *s, wr_film **out) {
    if (!s || !out) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    int rc = set_device(s->dev);
    if (rc) return rc;
    wr_film *f = new wr_film;
    f->scene = s;
    f->W = s->W;
    f->H = s->H;
    const size_t bytes = sizeof(double) * 3 * size_t(f->W) * f->H;
    hipError_t e = hipMalloc(&f->d_base, bytes);
    if (e == hipSuccess) e = hipMalloc(&f->d_splat, bytes);
    if (e != hipSuccess) {
        (void)hipFree(f->d_base);
        delete f;
        return set_error(e == hipErrorOutOfMemory ? WR_ERR_OUT_OF_MEMORY : WR_ERR_DEVICE, hipGetErrorString(e));
    }
    *out = f;
    return wr_film_clear(f);
}

void wr_film_destroy(wr_film *f) {
    if (!f) return;
    (void)hipSetDevice(f->scene->dev->ordinal);
    (void)hipStreamSynchronize(f->scene->stream);
    (void)hipFree(f->d_base);
    (void)hipFree(f->d_splat);
    delete f;
}

int wr_film_clear(wr_film *f) {
    if (!f) return set_error(WR_ERR_INVALID_ARGUMENT, "film is null");
    int rc = set_device(f->scene->dev);
    if (rc) return rc;
    const size_t n = 3 * size_t(f->W) * f->H;
    k_fill_f64<<<1024, 256, 0, f->scene->stream>>>(f->d_base, n);
    k_fill_f64<<<1024, 256, 0, f->scene->stream>>>(f->d_splat, n);
    WR_HIP(hipGetLastError());
    WR_HIP(hipStreamSynchronize(f->scene->stream));
    return WR_OK;
}

int wr_film_read(wr_film *f, double *base, double *splat) {
    if (!f) return set_error(WR_ERR_INVALID_ARGUMENT, "film is null");
    int rc = set_device(f->scene->dev);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * 3 * size_t(f->W) * f->H;
    WR_HIP(hipStreamSynchronize(f->scene->stream));
    if (base) WR_HIP(hipMemcpy(base, f->d_base, bytes, hipMemcpyDeviceToHost));
    if (splat) WR_HIP(hipMemcpy(splat, f->d_splat, bytes, hipMemcpyDeviceToHost));
    return WR_OK;
}

// Path pool: two queue-ordered states of 16*5 bytes per path (ray_o, ray_d,
// thr, rad, meta), the hit records (8 B), two shadow entries (2 * 48 B), and
// the per-sample radiance of a pass (16 B per sample).
// Byte layout of a path pool of `cap` paths and `samples` per-sample slots
// (every array 256-byte aligned): the state pair, hits, per-sample radiance,
// shadow queue (2 entries per path), counters (one 256-byte line each), stats.
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t pool_bytes(size_t cap, size_t samples) {
    return 10 * align256(16 * cap) + align256(8 * cap) + align256(16 * samples) + 3 * align256(32 * cap) +
           align256(4 * kCounterStride * kNumCounterLines) + align256(64);
}
// Points pb's arrays into the pool at `base`; returns the bytes used (pool_bytes).
static size_t pool_layout(char *base, uint32_t cap, uint32_t samples, PathBuffers *pb) {
    const size_t c = cap;
    size_t off = 0;
    auto take = [&](size_t bytes) { char *q = base + off; off += align256(bytes); return q; };
    PathState *sts[2] = {&pb->cur, &pb->next};
    for (PathState *st : sts) {
        st->ray_o = reinterpret_cast<float4 *>(take(16 * c));
        st->ray_d = reinterpret_cast<float4 *>(take(16 * c));
        st->thr = reinterpret_cast<float4 *>(take(16 * c));
        st->rad = reinterpret_cast<float4 *>(take(16 * c));
        st->meta = reinterpret_cast<uint4 *>(take(16 * c));
    }
    pb->hit = reinterpret_cast<uint2 *>(take(8 * c));
    pb->sample_rad = reinterpret_cast<float4 *>(take(16 * size_t(samples)));
    pb->sh_o = reinterpret_cast<float4 *>(take(32 * c));
    pb->sh_d = reinterpret_cast<float4 *>(take(32 * c));
    pb->sh_c = reinterpret_cast<float4 *>(take(32 * c));
    pb->counters = reinterpret_cast<uint32_t *>(take(4 * kCounterStride * kNumCounterLines));
    pb->cur.shadow_count = kShadowA;
    pb->next.shadow_count = kShadowB;
    pb->stats = reinterpret_cast<unsigned long long *>(take(64));
    return off;
}

static int ensure_capacity(wr_scene *s, uint32_t cap, uint32_t samples) {
    if (s->capacity >= cap && s->sample_capacity >= samples) return WR_OK;
    cap = std::max(cap, s->capacity);
    samples = std::max(samples, s->sample_capacity);
    if (s->pool) {
        WR_HIP(hipStreamSynchronize(s->stream));
        WR_HIP(hipFree(s->pool));
        s->pool = nullptr;
        s->capacity = 0;
        s->sample_capacity = 0;
    }
    char *p = nullptr;
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&p), pool_bytes(cap, samples)));
    s->pool = p;
    pool_layout(p, cap, samples, &s->pb);
    s->capacity = cap;
    s->sample_capacity = samples;
    if (!s->h_counters)
        WR_HIP(hipHostMalloc(reinterpret_cast<void **>(&s->h_counters), sizeof(uint32_t) * kCounterStride * kNumCounterLines,
                             hipHostMallocDefault));
    return WR_OK;
}

static int grid_for(const wr_scene *s, uint32_t n, int per_cu) {
    const uint32_t blocks = (n + kBlock - 1) / kBlock;
    return int(std::max<uint32_t>(1, std::min<uint32_t>(blocks, uint32_t(s->dev->n_cu) * per_cu)));
}
// Traversal grids index the spill region by block, so they never exceed max_trav_blocks.
static int trav_grid(const wr_scene *s, uint32_t n, int per_cu) {
    return std::min(grid_for(s, n, std::min(per_cu, kMaxTraversalBlocksPerCU)), s->max_trav_blocks);
}

// Samples of one pass: bounds the per-sample radiance buffer to 2 GiB, enough
// for a 1920x1080 frame at 64 spp, so such a frame drains the pool once.
constexpr uint64_t kMaxSamplesPerPass = uint64_t(1) << 27;
// Paths in flight when the caller leaves max_paths_in_flight at 0.
constexpr uint64_t kDefaultPool = uint64_t(1) << 27;
// The default pool of a scene whose pool has to grow: `want` paths, but no
// more than a quarter of the device memory that is free (counting the scene's
// current pool, which the growth releases) -- many scenes can then keep their
// pools in one process -- and never below the current pool or 2^20 paths.
static uint64_t default_pool_fit(const wr_scene *s, uint64_t want, uint64_t samples) {
    if (want <= s->capacity && samples <= s->sample_capacity) return want;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return want;
    uint64_t avail = free_b + (s->pool ? pool_bytes(s->capacity, s->sample_capacity) : 0);
    const uint64_t per_path = pool_bytes(1 << 20, 0) >> 20;
    const uint64_t budget = avail / 4 > 16 * samples ? (avail / 4 - 16 * samples) / per_path : 0;
    const uint64_t floor = std::min<uint64_t>(want, std::max<uint64_t>(s->capacity, uint64_t(1) << 20));
    return std::max(floor, std::min(want, budget));
}

// Runs wavefront iterations until the pool drains and both shadow queues are
// traced.  launch(pb) issues one iteration (refill, one traversal launch for
// the closest-hit rays and the previous shadow queue, shade, rotate) for the
// state pb; the host
// polls the device counters every kPoll iterations.  With a graph, the two
// iterations of a state round trip (cur -> next -> cur) are captured once and
// replayed, so a pass costs one graph launch per two iterations.  Every
// kernel reads its work counts from device memory, which is what makes the
// replays valid; nothing in an iteration is a memset.
template <class Launch>
static int drain(wr_scene *s, PathBuffers &pb, uint64_t max_iters, bool use_graph, uint32_t *host_counters,
                 Launch &&launch) {
    constexpr int kPoll = 4;
    auto poll = [&](bool &done) -> int {
        WR_HIP(hipMemcpyAsync(host_counters, pb.counters, sizeof(uint32_t) * kCounterStride * kNumCounterLines,
                              hipMemcpyDeviceToHost, s->stream));
        WR_HIP(hipStreamSynchronize(s->stream));
        // done: no path left and both shadow queues traced
        done = host_counters[kQ] == 0 && host_counters[kFresh] == 0 && host_counters[kShadowA] == 0 &&
               host_counters[kShadowB] == 0;
        return WR_OK;
    };
    if (!use_graph) {
        for (uint64_t it = 0;; it++) {
            if (it > max_iters) return set_error(WR_ERR_DEVICE, "path pool did not drain");
            launch(pb);
            WR_HIP(hipGetLastError());
            std::swap(pb.cur, pb.next);
            if (it % kPoll == kPoll - 1) {
                bool done = false;
                int rc = poll(done);
                if (rc) return rc;
                if (done) return WR_OK;
            }
        }
    }
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    WR_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    PathBuffers swapped = pb;
    std::swap(swapped.cur, swapped.next);
    launch(pb);
    launch(swapped);
    const hipError_t launch_err = hipGetLastError();
    WR_HIP(hipStreamEndCapture(s->stream, &graph));
    int rc = WR_OK;
    if (launch_err != hipSuccess) rc = set_error(WR_ERR_DEVICE, std::string("launch in graph capture: ") + hipGetErrorString(launch_err));
    if (!rc && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
        rc = set_error(WR_ERR_DEVICE, "hipGraphInstantiate failed");
    for (uint64_t it = 0; !rc; it += 2) {
        if (it > max_iters) {
            rc = set_error(WR_ERR_DEVICE, "path pool did not drain");
            break;
        }
        if (hipGraphLaunch(exec, s->stream) != hipSuccess) {
            rc = set_error(WR_ERR_DEVICE, "hipGraphLaunch failed");
            break;
        }
        if (it % kPoll == kPoll - 2) {
            bool done = false;
            rc = poll(done);
            if (done) break;
        }
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    return rc;  // an even number of iterations: pb is the current state again
}

// Largest traversal work chunk: every wave claims at least one chunk from a
// 32-bit work counter, so chunk x waves must stay far below 2^32 (a wrapped
// counter would hand out entries twice).
constexpr uint32_t kMaxFetchChunk = 4096;

int wr_render_params_validate(const wr_render_params *p) {
    if (!p) return set_error(WR_ERR_INVALID_ARGUMENT, "params is null");
    if (p->spp < 0) return set_error(WR_ERR_INVALID_ARGUMENT, "spp must be >= 0");
    if (p->first_sample < 0) return set_error(WR_ERR_INVALID_ARGUMENT, "first_sample must be >= 0");
    if (p->max_paths_in_flight < 0) return set_error(WR_ERR_INVALID_ARGUMENT, "max_paths_in_flight must be >= 0");
    if (p->fetch_chunk > kMaxFetchChunk)
        return set_error(WR_ERR_INVALID_ARGUMENT, "fetch_chunk must be 0 (default) or 1.." + std::to_string(kMaxFetchChunk));
    if (p->refill_lanes > 64) return set_error(WR_ERR_INVALID_ARGUMENT, "refill_lanes must be 0 (default) or 1..64");
    return WR_OK;
}

// Estimators of the frame loop.
enum Estimator { kEstNee, kEstPrimitive, kEstLight };
// A path pool and traversal spill region that the caller already holds (the
// allocation-free entry point's scratch): render_impl then allocates nothing.
struct FixedPool {
    PathBuffers pb;
    uint32_t capacity, sample_capacity;  // paths in flight, samples per pass
    uint2 *spill;                        // kSpillDepth entries per lane of trav_blocks blocks
    int trav_blocks;
};
// The frame loop of wr_render_nee, wr_render_guided (g non-null),
// wr_render_primitive, wr_render_light and wr_render_nee_noalloc (fixed).
static int render_impl(wr_scene *s, wr_film *f, const wr_render_params *prm, wr_guide *g, const wr_guided_params *gp,
                       wr_render_stats *stats, wr_guide_stats *gstats, Estimator est = kEstNee,
                       const FixedPool *fixed = nullptr) {
    if (!s || !f || !prm) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    if (f->scene != s) return set_error(WR_ERR_INVALID_ARGUMENT, "film belongs to another scene");
    if (int rc0 = wr_render_params_validate(prm)) return rc0;
    if (g && !gp) return set_error(WR_ERR_INVALID_ARGUMENT, "null guided parameters");
    if (g && guide_scene(g) != s) return set_error(WR_ERR_INVALID_ARGUMENT, "guide belongs to another scene");
    if (g && !(gp->guide_fraction >= 0.f && gp->guide_fraction < 1.f))
        return set_error(WR_ERR_INVALID_ARGUMENT, "guide_fraction must be in [0, 1)");
    const bool env = s->ds.env_select > 0.f;
    if (est != kEstPrimitive && s->ds.n_lights == 0 && !env) return set_error(WR_ERR_INVALID_ARGUMENT, "scene has no lights");
    if (env && est == kEstLight) return set_error(WR_ERR_UNSUPPORTED, "the light tracer does not support envmap lights");
    if (env && g) return set_error(WR_ERR_UNSUPPORTED, "guided rendering does not support envmap lights");
    if (env && s->ds.med.on) return set_error(WR_ERR_UNSUPPORTED, "an envmap with a medium is not supported");
    wr_guide_stats gst{};
    const auto t0 = std::chrono::steady_clock::now();
    wr_render_stats st{};
    int rc = set_device(s->dev);
    if (rc) return rc;
    const int N = prm->num_path_vertices;
    const uint32_t wh = uint32_t(s->W) * uint32_t(s->H);
    if (N <= 1 || prm->spp == 0) {  // pathnextevent.h:163
        if (stats) *stats = st;
        if (gstats) *gstats = gst;
        return WR_OK;
    }
    if (fixed) {
        // the pool is the caller's fixed region: nothing may grow, and nothing
        // that allocates (events for profiling, graph instantiation) may run
        if (prm->profile) return set_error(WR_ERR_UNSUPPORTED, "profile = 1 needs HIP events; the allocation-free entry point creates none");
        if (prm->use_graph) return set_error(WR_ERR_UNSUPPORTED, "use_graph = 1 instantiates a graph, which allocates");
        if (wh > fixed->sample_capacity)
            return set_error(WR_ERR_UNSUPPORTED, "frame of " + std::to_string(wh) + " pixels exceeds the allocation-free "
                             "pass capacity of " + std::to_string(fixed->sample_capacity) + " samples");
        if (uint64_t(prm->max_paths_in_flight) > fixed->capacity)
            return set_error(WR_ERR_UNSUPPORTED, "max_paths_in_flight exceeds the allocation-free pool of " +
                             std::to_string(fixed->capacity) + " paths");
    }
    uint64_t spp_cap = std::min<uint64_t>((fixed ? fixed->sample_capacity : kMaxSamplesPerPass) / wh, uint64_t(prm->spp));
    if (g && gp->spp_per_pass > 0) spp_cap = std::min<uint64_t>(spp_cap, gp->spp_per_pass);
    const uint32_t spp_pass = uint32_t(std::max<uint64_t>(1, spp_cap));
    const uint32_t samples_pass = spp_pass * wh;
    // default pool: 2^27 paths, a whole 1080p 64-spp pass (about 35 GB of state
    // at full size, capped by the pass): large launches hide the traversal
    // latency and shrink the per-launch tail (2^25 / 2^26 / 2^27 paths:
    // 5,090 / 5,308 / 5,442 Mrays/s on the 1.31M-triangle frame)
    uint64_t pool = prm->max_paths_in_flight > 0 ? uint64_t(prm->max_paths_in_flight) : kDefaultPool;
    if (fixed) pool = std::min<uint64_t>(pool, fixed->capacity);
    pool = std::min<uint64_t>(pool, samples_pass);
    if (!fixed && prm->max_paths_in_flight == 0) pool = default_pool_fit(s, pool, samples_pass);
    if (pool > (uint64_t(1) << 30)) return set_error(WR_ERR_UNSUPPORTED, "too many paths in flight");
    const uint32_t P = uint32_t(pool);
    uint32_t fixed_counters[kCounterStride * kNumCounterLines];  // host copy of the counters (fixed pool)
    uint32_t *host_counters = fixed_counters;
    uint2 *spill = s->d_spill;
    if (fixed) {
        spill = fixed->spill;
    } else {
        rc = ensure_capacity(s, P, samples_pass);
        if (rc) return rc;
        host_counters = s->h_counters;
    }
    if (g) {
        rc = guide_reserve_draws(g, P);
        if (!rc) rc = guide_render_begin(g);
        if (rc) return rc;
    }
    PathBuffers pb = fixed ? fixed->pb : s->pb;
    pb.film_base = f->d_base;
    pb.film_splat = f->d_splat;
    k_fill_u32<<<1, 64, 0, s->stream>>>(reinterpret_cast<uint32_t *>(pb.stats), 0u, 12);
    WR_HIP(hipGetLastError());
    const bool prof = prm->profile != 0;
    const bool count = prm->count_traversal != 0;
    // per-kernel profiling records events between launches, which a graph cannot hold
    const bool use_graph = prm->use_graph != 0 && !prof;
    TravTune tune;
    if (prm->fetch_chunk) tune.chunk = prm->fetch_chunk;
    if (prm->refill_lanes) tune.refill = prm->refill_lanes;
    std::vector<hipEvent_t> evs;
    struct Mark { int kind; size_t e0; };
    std::vector<Mark> marks;
    auto ev_record = [&]() -> size_t {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        (void)hipEventRecord(e, s->stream);
        evs.push_back(e);
        return evs.size() - 1;
    };
    auto begin = [&](int kind) { if (prof) marks.push_back({kind, ev_record()}); };
    auto end = [&]() { if (prof) ev_record(); };
    const int g_regen = std::max(1, std::min<int>((P + kShadeBlock - 1) / kShadeBlock, s->dev->n_cu * s->dev->occ_generate));
    // shading kernel: plain, or extended for guiding and/or a medium
    // the light tracer's sampler is not volumetric (lightsplatting.h:34): it sees vacuum
    DScene ds = s->ds;
    if (est == kEstLight) ds.med.on = 0;
    const bool medium = ds.med.on != 0;
    const void *shade_fn = est == kEstLight ? reinterpret_cast<const void *>(&k_shade_light)
                           : env            ? reinterpret_cast<const void *>(&k_shade<false, true>)
                                            : reinterpret_cast<const void *>(&k_shade<false>);
    if (g && medium) shade_fn = reinterpret_cast<const void *>(&k_shade_ext<true, true>);
    else if (g) shade_fn = reinterpret_cast<const void *>(&k_shade_ext<true, false>);
    else if (medium) shade_fn = reinterpret_cast<const void *>(&k_shade_ext<false, true>);
    const int g_shade = std::max(1, std::min<int>((P + kShadeBlock - 1) / kShadeBlock,
                                                  s->dev->n_cu * occupancy(s->dev, shade_fn, kShadeBlock)));
    int pass = 0;
    const bool fp8 = g && gp->fp8_inference != 0;
    const void *infer_fn = medium ? (fp8 ? reinterpret_cast<const void *>(&k_guide_infer<true, true>)
                                         : reinterpret_cast<const void *>(&k_guide_infer<true, false>))
                                  : (fp8 ? reinterpret_cast<const void *>(&k_guide_infer<false, true>)
                                         : reinterpret_cast<const void *>(&k_guide_infer<false, false>));
    const int g_infer = std::max(1, std::min<int>((P + kInferBlock - 1) / kInferBlock,
                                                  s->dev->n_cu * occupancy(s->dev, infer_fn, kInferBlock)));
    guide::Net net{};
    // up to P closest-hit rays and 2P shadow rays (one per path and one per new sample)
    int g_trav = trav_grid(s, 2 * P, s->dev->occ_rays);
    if (fixed) g_trav = std::min(g_trav, fixed->trav_blocks);  // the fixed spill region covers trav_blocks blocks
    for (uint32_t done = 0; done < uint32_t(prm->spp); done += spp_pass) {
        const uint32_t k = std::min<uint32_t>(spp_pass, uint32_t(prm->spp) - done);
        RenderArgs a;
        a.total = k * wh;
        a.wh = wh;
        a.first_sample = uint32_t(prm->first_sample) + done;
        a.seed = prm->seed;
        a.capacity = P;
        a.N = N;
        a.use_mis = est == kEstNee && prm->use_mis ? 1 : 0;
        a.nee = est == kEstNee ? 1 : 0;
        a.light = est == kEstLight ? 1 : 0;
        guide::Shade gs{};
        float record_prob = 0.f;
        const bool collect = g && gp->train_steps > 0;
        const int set = pass & 1;  // record set and inference copy of this pass
        if (g) {
            rc = guide_pass_net(g, set, gp->overlap_training != 0, &net);
            if (!rc && collect) rc = guide_begin_pass(g, set, a.total, &record_prob);
            if (rc) return rc;
            const bool sample_net = pass >= int(gp->warmup_passes);
            rc = guide_shade_view(g, set, sample_net ? gp->guide_fraction : 0.f, record_prob, collect,
                                  gp->max_guided_bounce ? gp->max_guided_bounce : 0xFFFFFFFFu, &gs);
            if (rc) return rc;
        }
        k_init<<<1, 64, 0, s->stream>>>(pb, a);
        WR_HIP(hipGetLastError());
        // Each iteration refills the pool, extends every active path by one
        // vertex and resolves the shadow rays of the previous iteration's
        // vertices (so the last shadow queue takes one iteration more);
        // iterations past the end find empty queues.  Every path ends within
        // N-1 vertices, and each iteration that is not draining starts new
        // samples in all free entries.
        const uint64_t max_iters = (uint64_t(a.total) / P + 2) * uint64_t(N + 1) + 64;
        auto iteration = [&](const PathBuffers &q) {
            begin(3);
            if (env) k_regen<false, true><<<g_regen, kShadeBlock, 0, s->stream>>>(ds, nullptr, 0, q, a);
            else k_regen<false><<<g_regen, kShadeBlock, 0, s->stream>>>(ds, nullptr, 0, q, a);
            end();
            begin(0);
            if (count) k_trace_bounce<true, false><<<g_trav, kBlock, 0, s->stream>>>(ds, nullptr, q, spill, tune);
            else k_trace_bounce_fast<<<g_trav, kBlock, 0, s->stream>>>(ds, q, spill, tune);
            end();
            begin(2);
            if (g && gs.alpha > 0.f) {
                if (medium && fp8)
                    k_guide_infer<true, true><<<g_infer, kInferBlock, 0, s->stream>>>(ds, q.cur, q.hit, q.counters, N,
                                                                                      net, gs.alpha, gs.max_bounce, gs.draw);
                else if (medium)
                    k_guide_infer<true, false><<<g_infer, kInferBlock, 0, s->stream>>>(ds, q.cur, q.hit, q.counters, N,
                                                                                       net, gs.alpha, gs.max_bounce, gs.draw);
                else if (fp8)
                    k_guide_infer<false, true><<<g_infer, kInferBlock, 0, s->stream>>>(ds, q.cur, q.hit, q.counters, N,
                                                                                       net, gs.alpha, gs.max_bounce, gs.draw);
                else
                    k_guide_infer<false, false><<<g_infer, kInferBlock, 0, s->stream>>>(ds, q.cur, q.hit, q.counters, N,
                                                                                        net, gs.alpha, gs.max_bounce, gs.draw);
            }
            if (g && medium) k_shade_ext<true, true><<<g_shade, kShadeBlock, 0, s->stream>>>(ds, q, a, gs);
            else if (g) k_shade_ext<true, false><<<g_shade, kShadeBlock, 0, s->stream>>>(ds, q, a, gs);
            else if (medium) k_shade_ext<false, true><<<g_shade, kShadeBlock, 0, s->stream>>>(ds, q, a, gs);
            else if (est == kEstLight) k_shade_light<<<g_shade, kShadeBlock, 0, s->stream>>>(ds, q, a);
            else if (env) k_shade<false, true><<<g_shade, kShadeBlock, 0, s->stream>>>(ds, nullptr, q, a);
            else k_shade<false><<<g_shade, kShadeBlock, 0, s->stream>>>(ds, nullptr, q, a);
            end();
            k_rotate<<<1, 1, 0, s->stream>>>(q, a);
        };
        rc = drain(s, pb, max_iters, use_graph, host_counters, iteration);
        if (rc) return rc;
        begin(4);
        k_accumulate<<<std::max<uint32_t>(1, std::min<uint32_t>((wh + 255) / 256, uint32_t(s->dev->n_cu) * 8)), 256, 0,
                       s->stream>>>(pb.sample_rad, pb.film_base, wh, k);
        end();
        WR_HIP(hipGetLastError());
        if (collect) {
            rc = guide_end_pass(g, set, pb.sample_rad, a.total, record_prob, gp->train_steps, &gst);
            if (rc) return rc;
        }
        pass++;
    }
    if (g) {
        rc = guide_render_end(g, &gst);
        if (rc) return rc;
    }
    unsigned long long hs[6] = {0};
    WR_HIP(hipMemcpyAsync(hs, pb.stats, sizeof hs, hipMemcpyDeviceToHost, s->stream));
    WR_HIP(hipStreamSynchronize(s->stream));
    st.paths = uint64_t(prm->spp) * wh;
    st.extend_rays = hs[0];
    st.shadow_rays = hs[1];
    st.extend_nodes = hs[2];
    st.extend_tris = hs[3];
    st.shadow_nodes = hs[4];
    st.shadow_tris = hs[5];
    if (prof) {
        for (const Mark &m : marks) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, evs[m.e0], evs[m.e0 + 1]);
            switch (m.kind) {
            case 0: st.extend_ms += ms; st.extend_launches++; break;
            case 1: st.shadow_ms += ms; st.shadow_launches++; break;
            case 2: st.shade_ms += ms; break;
            case 3: st.generate_ms += ms; break;
            default: st.accumulate_ms += ms; break;
            }
        }
    }
    for (hipEvent_t e : evs) (void)hipEventDestroy(e);
    st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    if (gstats) *gstats = gst;
    return WR_OK;
}

int wr_render_nee(wr_scene *s, wr_film *f, const wr_render_params *prm, wr_render_stats *stats) {
    return render_impl(s, f, prm, nullptr, nullptr, stats, nullptr);
}

// ============================================================== allocation-free NEE + MIS
// wr_render_nee_noalloc renders through a fixed region of device memory that
// the library reserves once per GPU instead of the scene's lazily grown path
// pool: a pool of 2^17 paths with 2^20 per-sample slots (passes of at most
// 2^20 samples) and the traversal spill of 1024 blocks -- 77 MiB, held by each
// GPU the library runs on until the process ends.  The region is a file-scope
// device array; its device address on each GPU comes from hipGetSymbolAddress
// at wr_device_create (the array's name in host code is not a device address).
constexpr uint32_t kFixedPaths = 1u << 17;
constexpr uint32_t kFixedSamples = 1u << 20;
constexpr int kFixedTravBlocks = 1024;
constexpr size_t kFixedPoolBytes = pool_bytes(kFixedPaths, kFixedSamples);
constexpr size_t kFixedSpillBytes = sizeof(uint2) * size_t(kFixedTravBlocks) * kBlock * kSpillDepth;
constexpr size_t kScratchBytes = kFixedPoolBytes + kFixedSpillBytes;
static_assert(kFixedPoolBytes % 256 == 0, "the spill region starts 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on the device current in wr_device_create.
static int scratch_setup(wr_device *d) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return set_error(WR_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    d->scratch = ptr;
    return WR_OK;
}

int wr_render_nee_noalloc(wr_scene *s, wr_film *f, const wr_render_params *prm, wr_render_stats *stats) {
    if (!s || !f || !prm) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    if (!s->dev->scratch) return set_error(WR_ERR_DEVICE, "the device has no scratch region");
    char *base = static_cast<char *>(s->dev->scratch);
    FixedPool fp{};
    pool_layout(base, kFixedPaths, kFixedSamples, &fp.pb);
    fp.capacity = kFixedPaths;
    fp.sample_capacity = kFixedSamples;
    fp.spill = reinterpret_cast<uint2 *>(base + kFixedPoolBytes);
    fp.trav_blocks = kFixedTravBlocks;
    return render_impl(s, f, prm, nullptr, nullptr, stats, nullptr, kEstNee, &fp);
}

int wr_render_guided(wr_scene *s, wr_film *f, wr_guide *g, const wr_render_params *prm, const wr_guided_params *gp,
                     wr_render_stats *stats, wr_guide_stats *gstats) {
    if (!g) return set_error(WR_ERR_INVALID_ARGUMENT, "guide is null");
    return render_impl(s, f, prm, g, gp, stats, gstats);
}

int wr_render_primitive(wr_scene *s, wr_film *f, const wr_render_params *prm, wr_render_stats *stats) {
    return render_impl(s, f, prm, nullptr, nullptr, stats, nullptr, kEstPrimitive);
}

int wr_render_light(wr_scene *s, wr_film *f, const wr_render_params *prm, wr_render_stats *stats) {
    return render_impl(s, f, prm, nullptr, nullptr, stats, nullptr, kEstLight);
}

int wr_render_visualize(wr_scene *s, wr_film *f, const wr_render_params *prm, wr_render_stats *stats) {
    if (!s || !f || !prm) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    if (f->scene != s) return set_error(WR_ERR_INVALID_ARGUMENT, "film belongs to another scene");
    if (int rc0 = wr_render_params_validate(prm)) return rc0;
    const auto t0 = std::chrono::steady_clock::now();
    int rc = set_device(s->dev);
    if (rc) return rc;
    wr_render_stats st{};
    const uint32_t wh = uint32_t(s->W) * uint32_t(s->H);
    if (prm->spp > 0) {
        const int grid = trav_grid(s, wh, s->dev->occ_trace);
        if (prm->profile) WR_HIP(hipEventRecord(s->ev[0], s->stream));
        k_visualize<<<grid, kBlock, 0, s->stream>>>(s->ds, wh, uint32_t(prm->spp), uint32_t(prm->first_sample),
                                                    prm->seed, f->d_base, s->d_spill);
        WR_HIP(hipGetLastError());
        if (prm->profile) WR_HIP(hipEventRecord(s->ev[1], s->stream));
        WR_HIP(hipStreamSynchronize(s->stream));
        if (prm->profile) {
            float ms = 0.f;
            WR_HIP(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
            st.extend_ms = ms;
            st.extend_launches = 1;
        }
    }
    st.paths = uint64_t(prm->spp) * wh;
    st.extend_rays = st.paths;
    st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    return WR_OK;
}

int wr_render_bdpt(wr_scene *s, wr_film *f, const wr_render_params *prm, wr_render_stats *stats) {
    if (!s || !f || !prm) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    if (f->scene != s) return set_error(WR_ERR_INVALID_ARGUMENT, "film belongs to another scene");
    if (int rc0 = wr_render_params_validate(prm)) return rc0;
    const auto t0 = std::chrono::steady_clock::now();
    wr_render_stats st{};
    const int N = prm->num_path_vertices;
    if (N <= 1 || prm->spp == 0) {  // BidirPathRenderer::render returns at once for N <= 1 (bidirpath.h:318)
        st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (stats) *stats = st;
        return WR_OK;
    }
    if (N > kBdptMaxVertices) return set_error(WR_ERR_UNSUPPORTED, "bidirectional path tracing supports at most 64 path vertices");
    if (s->ds.n_lights == 0) return set_error(WR_ERR_INVALID_ARGUMENT, "scene has no lights");
    if (s->ds.env_select > 0.f) return set_error(WR_ERR_UNSUPPORTED, "the bidirectional path tracer does not support envmap lights");
    int rc = set_device(s->dev);
    if (rc) return rc;
    const uint32_t wh = uint32_t(s->W) * uint32_t(s->H);
    const int grid = trav_grid(s, wh, occupancy(s->dev, reinterpret_cast<const void *>(&k_bdpt), kBlock));
    // records of N - 1 light vertices per lane, behind 256 bytes of counters
    const size_t need = 256 + sizeof(float4) * 4 * size_t(N - 1) * size_t(grid) * kBlock;
    if (need > s->bdpt_bytes) {
        WR_HIP(hipStreamSynchronize(s->stream));
        (void)hipFree(s->d_bdpt);
        s->device_bytes -= s->bdpt_bytes;
        s->d_bdpt = nullptr;
        s->bdpt_bytes = 0;
        WR_HIP(hipMalloc(&s->d_bdpt, need));
        s->bdpt_bytes = need;
        s->device_bytes += need;
    }
    auto *counts = reinterpret_cast<unsigned long long *>(s->d_bdpt);
    WR_HIP(hipMemsetAsync(counts, 0, 16, s->stream));
    // the medium is not passed on: the reference's subpath sampler here is not volumetric
    DScene ds = s->ds;
    ds.med.on = 0;
    if (prm->profile) WR_HIP(hipEventRecord(s->ev[0], s->stream));
    k_bdpt<<<grid, kBlock, 0, s->stream>>>(ds, wh, uint32_t(prm->spp), uint32_t(prm->first_sample), prm->seed, N,
                                           f->d_base, f->d_splat, s->d_bdpt + 16, counts, s->d_spill);
    WR_HIP(hipGetLastError());
    if (prm->profile) WR_HIP(hipEventRecord(s->ev[1], s->stream));
    unsigned long long h[2] = {0, 0};
    WR_HIP(hipMemcpyAsync(h, counts, sizeof h, hipMemcpyDeviceToHost, s->stream));
    WR_HIP(hipStreamSynchronize(s->stream));
    if (prm->profile) {
        float ms = 0.f;
        WR_HIP(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
        st.extend_ms = ms;
        st.extend_launches = 1;
    }
    st.paths = uint64_t(prm->spp) * wh;
    st.extend_rays = h[0];
    st.shadow_rays = h[1];
    st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    return WR_OK;
}

// ============================================================== batches (host)
int wr_render_batch(const wr_batch_job *jobs, uint32_t n_jobs, const wr_render_params *prm, wr_render_stats *stats) {
    if (!jobs || !prm) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    if (int rc0 = wr_render_params_validate(prm)) return rc0;
    if (n_jobs == 0) return set_error(WR_ERR_INVALID_ARGUMENT, "empty batch");
    if (n_jobs > uint32_t(kMaxBatchJobs)) return set_error(WR_ERR_UNSUPPORTED, "too many jobs in a batch");
    wr_scene *s0 = jobs[0].scene;
    for (uint32_t j = 0; j < n_jobs; j++) {
        const wr_batch_job &J = jobs[j];
        if (!J.scene || !J.film) return set_error(WR_ERR_INVALID_ARGUMENT, "batch job without scene or film");
        if (J.film->scene != J.scene) return set_error(WR_ERR_INVALID_ARGUMENT, "batch job film belongs to another scene");
        if (J.scene->dev != s0->dev) return set_error(WR_ERR_INVALID_ARGUMENT, "batch scenes on different devices");
        if (J.spp < 0 || J.first_sample < 0) return set_error(WR_ERR_INVALID_ARGUMENT, "negative spp or first_sample");
        if (J.num_path_vertices > int(kBounceMask)) return set_error(WR_ERR_UNSUPPORTED, "num_path_vertices above 255");
        if (uint64_t(J.scene->W) * J.scene->H > kMaxSamplesPerPass)
            return set_error(WR_ERR_UNSUPPORTED, "resolution too large for a batch pass");
        if (J.scene->ds.med.on) return set_error(WR_ERR_UNSUPPORTED, "batch scenes with a medium are not supported");
        if (J.scene->ds.n_lights == 0 && !(J.scene->ds.env_select > 0.f))
            return set_error(WR_ERR_INVALID_ARGUMENT, "scene has no lights");
    }
    const auto t0 = std::chrono::steady_clock::now();
    wr_render_stats st{};
    int rc = set_device(s0->dev);
    if (rc) return rc;
    std::vector<uint32_t> done(n_jobs, 0);
    uint64_t paths = 0, max_pass = 0;
    int maxN = 2;
    for (uint32_t j = 0; j < n_jobs; j++) {
        const uint64_t wh = uint64_t(jobs[j].scene->W) * jobs[j].scene->H;
        if (jobs[j].num_path_vertices > 1) {
            paths += wh * uint64_t(jobs[j].spp);
            maxN = std::max(maxN, jobs[j].num_path_vertices);
        }
        max_pass += wh * uint64_t(jobs[j].spp);
    }
    if (paths == 0) {
        if (stats) *stats = st;
        return WR_OK;
    }
    max_pass = std::min<uint64_t>(max_pass, kMaxSamplesPerPass);
    uint64_t pool = prm->max_paths_in_flight > 0 ? uint64_t(prm->max_paths_in_flight) : kDefaultPool;
    pool = std::min<uint64_t>(pool, max_pass);
    if (prm->max_paths_in_flight == 0) pool = default_pool_fit(s0, pool, max_pass);
    if (pool > (uint64_t(1) << 30)) return set_error(WR_ERR_UNSUPPORTED, "too many paths in flight");
    const uint32_t P = uint32_t(pool);
    rc = ensure_capacity(s0, P, uint32_t(max_pass));
    if (rc) return rc;
    PathBuffers pb = s0->pb;
    pb.film_base = nullptr;  // per job, from the table
    pb.film_splat = nullptr;
    BatchJob *d_tab = nullptr;
    WR_HIP(hipMalloc(reinterpret_cast<void **>(&d_tab), sizeof(BatchJob) * n_jobs));
    struct Free { BatchJob *p; ~Free() { if (p) (void)hipFree(p); } } free_tab{d_tab};
    k_fill_u32<<<1, 64, 0, s0->stream>>>(reinterpret_cast<uint32_t *>(pb.stats), 0u, 12);
    WR_HIP(hipGetLastError());
    const bool prof = prm->profile != 0;
    const bool count = prm->count_traversal != 0;
    const bool use_graph = prm->use_graph != 0 && !prof;
    TravTune tune;
    if (prm->fetch_chunk) tune.chunk = prm->fetch_chunk;
    if (prm->refill_lanes) tune.refill = prm->refill_lanes;
    std::vector<hipEvent_t> evs;
    struct Mark { int kind; size_t e0; };
    std::vector<Mark> marks;
    auto ev_record = [&]() -> size_t {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        (void)hipEventRecord(e, s0->stream);
        evs.push_back(e);
        return evs.size() - 1;
    };
    auto begin = [&](int kind) { if (prof) marks.push_back({kind, ev_record()}); };
    auto end = [&]() { if (prof) ev_record(); };
    wr_device *dev = s0->dev;
    auto grid = [&](const void *fn, int block) {
        return std::max(1, std::min<int>((P + block - 1) / block, dev->n_cu * occupancy(dev, fn, block)));
    };
    bool env = false;
    for (uint32_t j = 0; j < n_jobs; j++) env = env || jobs[j].scene->ds.env_select > 0.f;
    const int g_regen = grid(env ? reinterpret_cast<const void *>(&k_regen<true, true>)
                                 : reinterpret_cast<const void *>(&k_regen<true>), kShadeBlock);
    const int g_shade = grid(env ? reinterpret_cast<const void *>(&k_shade<true, true>)
                                 : reinterpret_cast<const void *>(&k_shade<true>), kShadeBlock);
    const int g_trav =
        trav_grid(s0, 2 * P, occupancy(dev, reinterpret_cast<const void *>(&k_trace_bounce<false, true>), kBlock));
    std::vector<BatchJob> tab;
    for (;;) {
        // plan the pass: whole spp planes of each unfinished job, job-major, within the pass budget
        tab.clear();
        std::vector<uint32_t> tab_job;
        uint64_t budget = kMaxSamplesPerPass, offset = 0;
        for (uint32_t j = 0; j < n_jobs; j++) {
            const wr_batch_job &J = jobs[j];
            if (J.num_path_vertices <= 1 || done[j] >= uint32_t(J.spp)) continue;
            const uint32_t wh = uint32_t(J.scene->W) * uint32_t(J.scene->H);
            const uint32_t k = uint32_t(std::min<uint64_t>(uint64_t(J.spp) - done[j], budget / wh));
            if (k == 0) continue;
            BatchJob b{};
            b.s = J.scene->ds;
            b.offset = uint32_t(offset);
            b.count = k * wh;
            b.wh = wh;
            b.k = k;
            b.first_sample = uint32_t(J.first_sample) + done[j];
            b.seed = J.seed;
            b.N = J.num_path_vertices;
            b.use_mis = J.use_mis ? 1 : 0;
            b.film_base = J.film->d_base;
            b.film_splat = J.film->d_splat;
            tab.push_back(b);
            tab_job.push_back(j);
            offset += b.count;
            budget -= b.count;
        }
        if (tab.empty()) break;
        WR_HIP(hipMemcpyAsync(d_tab, tab.data(), sizeof(BatchJob) * tab.size(), hipMemcpyHostToDevice, s0->stream));
        RenderArgs a{};
        a.total = uint32_t(offset);
        a.capacity = P;
        a.N = maxN;
        a.nee = 1;
        const uint32_t nt = uint32_t(tab.size());
        k_init<<<1, 64, 0, s0->stream>>>(pb, a);
        WR_HIP(hipGetLastError());
        const uint64_t max_iters = (uint64_t(a.total) / P + 2) * uint64_t(maxN + 1) + 64;
        auto iteration = [&](const PathBuffers &q) {
            begin(3);
            if (env) k_regen<true, true><<<g_regen, kShadeBlock, 0, s0->stream>>>(s0->ds, d_tab, nt, q, a);
            else k_regen<true><<<g_regen, kShadeBlock, 0, s0->stream>>>(s0->ds, d_tab, nt, q, a);
            end();
            begin(0);
            if (count) k_trace_bounce<true, true><<<g_trav, kBlock, 0, s0->stream>>>(s0->ds, d_tab, q, s0->d_spill, tune);
            else k_trace_bounce<false, true><<<g_trav, kBlock, 0, s0->stream>>>(s0->ds, d_tab, q, s0->d_spill, tune);
            end();
            begin(2);
            if (env) k_shade<true, true><<<g_shade, kShadeBlock, 0, s0->stream>>>(s0->ds, d_tab, q, a);
            else k_shade<true><<<g_shade, kShadeBlock, 0, s0->stream>>>(s0->ds, d_tab, q, a);
            end();
            k_rotate<<<1, 1, 0, s0->stream>>>(q, a);
        };
        rc = drain(s0, pb, max_iters, use_graph, s0->h_counters, iteration);
        if (rc) return rc;
        begin(4);
        for (uint32_t t = 0; t < nt; t++) {
            const BatchJob &b = tab[t];
            k_accumulate<<<std::max<uint32_t>(1, std::min<uint32_t>((b.wh + 255) / 256, uint32_t(dev->n_cu) * 8)), 256, 0,
                           s0->stream>>>(pb.sample_rad + b.offset, b.film_base, b.wh, b.k);
            done[tab_job[t]] += b.k;
        }
        end();
        WR_HIP(hipGetLastError());
    }
    unsigned long long hs[6] = {0};
    WR_HIP(hipMemcpyAsync(hs, pb.stats, sizeof hs, hipMemcpyDeviceToHost, s0->stream));
    WR_HIP(hipStreamSynchronize(s0->stream));
    st.paths = paths;
    st.extend_rays = hs[0];
    st.shadow_rays = hs[1];
    st.extend_nodes = hs[2];
    st.extend_tris = hs[3];
    st.shadow_nodes = hs[4];
    st.shadow_tris = hs[5];
    for (const Mark &m : marks) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, evs[m.e0], evs[m.e0 + 1]);
        switch (m.kind) {
        case 0: st.extend_ms += ms; st.extend_launches++; break;
        case 1: st.shadow_ms += ms; st.shadow_launches++; break;
        case 2: st.shade_ms += ms; break;
        case 3: st.generate_ms += ms; break;
        default: st.accumulate_ms += ms; break;
        }
    }
    for (hipEvent_t e : evs) (void)hipEventDestroy(e);
    st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    return WR_OK;
}

int wr_autotune(const wr_batch_job *jobs, uint32_t n_jobs, const wr_render_params *base, wr_render_params *best,
                double *mrays) {
    if (!jobs || !base || !best) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    if (int rc0 = wr_render_params_validate(base)) return rc0;
    if (n_jobs == 0) return set_error(WR_ERR_INVALID_ARGUMENT, "empty batch");
    // trial jobs: 1 spp each into scratch films
    std::vector<wr_batch_job> trial(jobs, jobs + n_jobs);
    std::vector<wr_film *> films;
    int rc = WR_OK;
    for (auto &t : trial) {
        if (!t.scene) { rc = set_error(WR_ERR_INVALID_ARGUMENT, "batch job without scene"); break; }
        wr_film *f = nullptr;
        rc = wr_film_create(t.scene, &f);
        if (rc) break;
        films.push_back(f);
        t.film = f;
        t.spp = std::min(t.spp, 1);
    }
    struct Candidate { uint32_t pool, chunk, refill; int graph; };
    std::vector<Candidate> cands;
    for (uint32_t pool : {1u << 23, 1u << 24, 1u << 25, 1u << 26, 1u << 27})
        for (auto ck : {std::make_pair(96u, 16u), std::make_pair(256u, 16u), std::make_pair(128u, 32u), std::make_pair(256u, 32u)})
            for (int graph : {0, 1}) cands.push_back({pool, ck.first, ck.second, graph});
    double best_rate = -1.0;
    wr_render_params bp = *base;
    if (!rc) {
        wr_render_params warm = *base;
        warm.profile = 0;
        warm.count_traversal = 0;
        wr_render_stats ws;
        rc = wr_render_batch(trial.data(), n_jobs, &warm, &ws);  // allocations, first launches
    }
    for (size_t c = 0; !rc && c < cands.size(); c++) {
        wr_render_params p = *base;
        p.profile = 0;
        p.count_traversal = 0;
        p.max_paths_in_flight = int32_t(cands[c].pool);
        p.fetch_chunk = cands[c].chunk;
        p.refill_lanes = cands[c].refill;
        p.use_graph = cands[c].graph;
        wr_render_stats st;
        rc = wr_render_batch(trial.data(), n_jobs, &p, &st);
        if (rc) break;
        const double rate = double(st.extend_rays + st.shadow_rays) / (st.wall_ms * 1e-3) / 1e6;
        if (rate > best_rate) {
            best_rate = rate;
            bp = p;
        }
    }
    const std::string err = rc ? wr::error_string() : std::string();
    for (wr_film *f : films) wr_film_destroy(f);
    if (rc) {
        wr::error_string() = err;
        return rc;
    }
    bp.profile = base->profile;
    bp.count_traversal = base->count_traversal;
    *best = bp;
    if (mrays) *mrays = best_rate;
    return WR_OK;
}

static int trace_impl(wr_scene *s, const float *d_rays, uint32_t n, void *d_hits, uint8_t *d_occ, bool any, float *ms) {
    int rc = set_device(s->dev);
    if (rc) return rc;
    if (n == 0) { if (ms) *ms = 0.f; return WR_OK; }
    const int grid = trav_grid(s, n, s->dev->occ_trace);
    if (ms) WR_HIP(hipEventRecord(s->ev[0], s->stream));
    if (any) k_trace<true><<<grid, kBlock, 0, s->stream>>>(s->ds, reinterpret_cast<const float4 *>(d_rays), n, nullptr,
                                                           d_occ, s->d_spill);
    else k_trace<false><<<grid, kBlock, 0, s->stream>>>(s->ds, reinterpret_cast<const float4 *>(d_rays), n,
                                                         reinterpret_cast<uint2 *>(d_hits), nullptr, s->d_spill);
    WR_HIP(hipGetLastError());
    if (ms) {
        WR_HIP(hipEventRecord(s->ev[1], s->stream));
        WR_HIP(hipEventSynchronize(s->ev[1]));
        WR_HIP(hipEventElapsedTime(ms, s->ev[0], s->ev[1]));
    }
    return WR_OK;
}

int wr_trace_closest_device(wr_scene *s, const float *d_rays, uint32_t n, void *d_hits, float *ms) {
    if (!s || (n && (!d_rays || !d_hits))) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    return trace_impl(s, d_rays, n, d_hits, nullptr, false, ms);
}

int wr_trace_any_device(wr_scene *s, const float *d_rays, uint32_t n, uint8_t *d_occ, float *ms) {
    if (!s || (n && (!d_rays || !d_occ))) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    return trace_impl(s, d_rays, n, nullptr, d_occ, true, ms);
}

int wr_trace_closest(wr_scene *s, const float *rays, uint32_t n, uint32_t *geom, uint32_t *prim, float *t) {
    if (!s || (n && (!rays || !geom || !prim || !t))) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0) return WR_OK;
    int rc = set_device(s->dev);
    if (rc) return rc;
    float *d_rays = nullptr;
    uint2 *d_hits = nullptr;
    WR_HIP(hipMalloc(&d_rays, sizeof(float) * 8 * size_t(n)));
    hipError_t e = hipMalloc(&d_hits, sizeof(uint2) * size_t(n));
    if (e != hipSuccess) { (void)hipFree(d_rays); return set_error(WR_ERR_OUT_OF_MEMORY, hipGetErrorString(e)); }
    std::vector<uint2> hits(n);
    e = hipMemcpy(d_rays, rays, sizeof(float) * 8 * size_t(n), hipMemcpyHostToDevice);
    if (e == hipSuccess) rc = trace_impl(s, d_rays, n, d_hits, nullptr, false, nullptr);
    if (e == hipSuccess && rc == WR_OK) e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess && rc == WR_OK) e = hipMemcpy(hits.data(), d_hits, sizeof(uint2) * size_t(n), hipMemcpyDeviceToHost);
    (void)hipFree(d_rays);
    (void)hipFree(d_hits);
    if (rc) return rc;
    if (e != hipSuccess) return set_error(WR_ERR_DEVICE, hipGetErrorString(e));
    for (uint32_t i = 0; i < n; i++) {
        if (hits[i].x == kInvalid) {
            geom[i] = prim[i] = WR_INVALID_ID;
            t[i] = INFINITY;
        } else {
            const uint32_t orig = s->leaf_to_orig[hits[i].x];
            geom[i] = s->tri_geom[orig];
            prim[i] = orig - s->geom_first[geom[i]];
            std::memcpy(&t[i], &hits[i].y, 4);
        }
    }
    return WR_OK;
}

int wr_trace_any(wr_scene *s, const float *rays, uint32_t n, uint8_t *occluded) {
    if (!s || (n && (!rays || !occluded))) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0) return WR_OK;
    int rc = set_device(s->dev);
    if (rc) return rc;
    float *d_rays = nullptr;
    uint8_t *d_occ = nullptr;
    WR_HIP(hipMalloc(&d_rays, sizeof(float) * 8 * size_t(n)));
    hipError_t e = hipMalloc(&d_occ, size_t(n));
    if (e != hipSuccess) { (void)hipFree(d_rays); return set_error(WR_ERR_OUT_OF_MEMORY, hipGetErrorString(e)); }
    e = hipMemcpy(d_rays, rays, sizeof(float) * 8 * size_t(n), hipMemcpyHostToDevice);
    if (e == hipSuccess) rc = trace_impl(s, d_rays, n, nullptr, d_occ, true, nullptr);
    if (e == hipSuccess && rc == WR_OK) e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess && rc == WR_OK) e = hipMemcpy(occluded, d_occ, size_t(n), hipMemcpyDeviceToHost);
    (void)hipFree(d_rays);
    (void)hipFree(d_occ);
    if (rc) return rc;
    if (e != hipSuccess) return set_error(WR_ERR_DEVICE, hipGetErrorString(e));
    return WR_OK;
}

