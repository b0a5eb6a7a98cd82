// wr_host.h -- host-side state shared by the library's translation units:
// error reporting, the path-pool layout and the device, scene and film handles.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/wurst_amd.h"
#include "wr_device.h"
#include "wr_layout.h"

namespace wr {
// Thread-local message behind wr_last_error().
inline std::string &error_string() {
    static thread_local std::string s;
    return s;
}
}  // namespace wr

using namespace wr;

// ============================================================== errors
static inline int set_error(int code, const std::string &msg) {
    wr::error_string() = msg;
    return code;
}
#define WR_HIP(expr)                                                                                     \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return set_error(e_ == hipErrorOutOfMemory ? WR_ERR_OUT_OF_MEMORY : WR_ERR_DEVICE,           \
                             std::string(#expr) + ": " + hipGetErrorString(e_));                         \
    } while (0)

// ============================================================== kernel arguments
constexpr int kBlock = 64;  // one wavefront per block; LDS stack indexed by lane

// Shadow-ray targets: where a visible connection's contribution goes.
constexpr uint32_t kTargetMask = 0x3FFFFFFFu;
constexpr uint32_t kToNextPath = 0u << 30;   // radiance of a continuing path (its index in the state that
                                             // the shade kernel wrote, current when the ray is traced)
constexpr uint32_t kToSample = 1u << 30;     // per-sample radiance of a finished path
constexpr uint32_t kToSplat = 2u << 30;      // light->camera splat (pixel of the FP64 splat film)

// Path state in queue order (SoA, double buffered): entry i of the current
// state is the i-th active path.  Shade writes continuing paths compacted into
// the next state; regen appends new samples behind them.  Every kernel walks
// the state in order, so all path-state traffic is coalesced.
struct PathState {
    float4 *ray_o, *ray_d;  // ray to extend
    float4 *thr;            // rgb: throughput of the next vertex; w: pdfW of the ray direction
    float4 *rad;            // rgb: radiance gathered so far (FP32)
    uint4 *meta;            // x: RNG key, y: sample index in the pass, z: bounce of the next vertex,
                            // w: 1 + training record of the previous vertex (0: none)
    uint32_t shadow_count;  // counter (kShadowA or kShadowB) of the shadow queue that the shade
                            // kernel appends to while this state is current; the two states
                            // hold the two counters and swap them with everything else
};

struct PathBuffers {
    PathState cur, next;
    uint2 *hit;                  // per current entry: leaf-order triangle (or kInvalid), t bits
    float4 *sample_rad;          // per sample of the pass: final radiance (FP32)
    float4 *sh_o, *sh_d, *sh_c;  // shadow rays: o.xyz+tmax, d.xyz+target, contribution
    uint32_t *counters;          // see Counter
    unsigned long long *stats;   // [0] extend rays [1] shadow rays [2] ext nodes [3] ext tris [4] shd nodes [5] shd tris
    double *film_base, *film_splat;
};

// Each counter sits on its own 256-byte line: atomics on one line serialise.
constexpr int kCounterStride = 64;
constexpr int kXcds = 8;  // MI355X: 8 accelerator complex dies, each with its own L2
enum Counter : int {
    kQ = 0 * kCounterStride,         // continuing paths at the head of the current state
    kFresh = 1 * kCounterStride,     // new samples started behind them this iteration
    kQn = 2 * kCounterStride,        // paths continuing into the next state
    // line 3 is spare
    // Shadow queue: the shade kernel of iteration i appends its shadow rays,
    // then the regeneration of iteration i + 1 its t = 1 connections, and the
    // traversal launch of iteration i + 1 traces them beside that iteration's
    // closest-hit rays.  Consecutive iterations alternate between the two
    // counters (PathState::shadow_count), so a queue's count survives the end
    // of its iteration while the next shade kernel appends from zero.
    kShadowA = 4 * kCounterStride,
    kShadowB = 5 * kCounterStride,
    kClaimed = 6 * kCounterStride,   // samples started so far in the pass
    kBase = 7 * kCounterStride,      // first sample index started this iteration
    kFetchExtX = 8 * kCounterStride,   // extend: work counter of XCD x's share at kFetchExtX + x * kCounterStride
    kFetchShdX = 16 * kCounterStride,  // shadow: likewise
    kNumCounterLines = 24
};

struct RenderArgs {
    uint32_t total;           // samples of this pass = spp * W * H
    uint32_t wh;
    uint32_t first_sample;    // sample index of the pass's first spp
    uint32_t seed;
    uint32_t capacity;        // path pool size
    int N;                    // num_path_vertices
    int use_mis;
    int nee;                  // 1: PathNeeRenderer (NEE, t = 1 splats); 0: PrimitivePathRenderer (BSDF sampling only)
    int light;                // 1: LightSplattingRenderer (light subpaths connected to the camera)
};

// ============================================================== host objects
struct wr_device {
    int ordinal = 0;
    int n_cu = 256;
    // resident blocks per CU of each kernel (HIP occupancy API), used to size persistent grids
    int occ_rays = 8, occ_shade = 8, occ_generate = 8, occ_trace = 8;
    std::map<const void *, int> occ_cache;  // other kernels, filled on first use
    void *scratch = nullptr;  // device address of wr_render_nee_noalloc's fixed region on this GPU
};

struct wr_scene {
    wr_device *dev = nullptr;
    hipStream_t stream = nullptr;
    DScene ds{};
    // device allocations
    Node4 *d_nodes = nullptr;   // the build's BVH4 (checks and statistics)
    NodeQ *d_qnodes = nullptr;  // its quantised copy, which the kernels traverse
    float4 *d_tris = nullptr, *d_ltris = nullptr;
    uint2 *d_spill = nullptr;   // traversal stack spill: kSpillDepth entries per lane of the largest grid
    float4 *d_bdpt = nullptr;   // bidirectional light-vertex records (+ 2 ray counters), grown on demand
    size_t bdpt_bytes = 0;
    int max_trav_blocks = 0;
    int2 *d_geom = nullptr;
    DMat *d_mats = nullptr;
    DLight *d_lights = nullptr;
    float4 *d_env = nullptr;            // envmap texels; cdf and pdf follow in d_env_cdf
    float *d_env_cdf = nullptr;
    std::vector<DLight> h_lights;       // host copy of the light table
    std::vector<double> light_w;        // area-light selection weights |L|_2
    uint64_t n_nodes = 0, n_tris = 0, device_bytes = 0;
    uint32_t max_stack = 0;
    int builder = WR_BVH_GPU;           // which builder made the BVH
    int ploc_radius = 0;                // its clustering window (GPU builder; 0 = default)
    uint32_t bvh_depth = 0, bvh_iterations = 0;
    double build_ms = 0.0;
    // host side: triangle index -> (geom, prim)
    std::vector<uint32_t> tri_geom, geom_first, leaf_to_orig;
    int W = 0, H = 0;
    float box[6] = {0, 0, 0, 0, 0, 0};  // bounds of the vertices (lo xyz, hi xyz)
    // path buffers (lazily sized)
    PathBuffers pb{};
    uint32_t capacity = 0, sample_capacity = 0;
    void *pool = nullptr;
    uint32_t *h_counters = nullptr;  // pinned host copy of the pool counters
    hipEvent_t ev[2] = {nullptr, nullptr};
};

struct wr_film {
    wr_scene *scene = nullptr;
    double *d_base = nullptr, *d_splat = nullptr;
    int W = 0, H = 0;
};

static inline int set_device(wr_device *dev) {
    WR_HIP(hipSetDevice(dev->ordinal));
    return WR_OK;
}

// ============================================================== GPU BVH build (wr_bvh_gpu.hip)
namespace wr {
struct GpuBvhResult {
    Node4 *nodes = nullptr;     // device, n_nodes, breadth-first (root = 0, children after parents)
    TriRecord *tris = nullptr;  // device, leaf order
    uint32_t n_nodes = 0, max_stack = 0, depth = 0, iterations = 0;
    double ms = 0.0;            // wall time of the build
};
// Builds from a mesh already in device memory (verts [3n_v], tris [3n], tri_geom [n],
// indices validated) with a clustering window of `radius` clusters on each side
// (0: default, at most 64); leaf_to_orig (host, n entries) receives the input
// triangle of each leaf-order record.  On error the caller frees whatever
// out->nodes / out->tris hold.
int build_bvh4_gpu(const float *d_verts, const uint32_t *d_tris, const uint32_t *d_tri_geom, uint32_t n,
                   hipStream_t stream, int radius, uint32_t *leaf_to_orig, GpuBvhResult *out);
}  // namespace wr

// ============================================================== guiding (wr_guide.hip)
#include "wr_guide.h"
// A render call: guide_render_begin; per pass (record set k = pass % 2):
// guide_pass_net (the network inference reads: the one trained on the previous
// pass; the render stream waits for that training on the GPU),
// guide_begin_pass, guide_shade_view, the pass, guide_end_pass (training
// weights from the pass's final per-sample radiance, then `steps` Adam steps
// enqueued on the guide's own stream); guide_render_end waits for the last
// training.
int guide_render_begin(wr_guide *g);
int guide_pass_net(wr_guide *g, int set, bool overlap, wr::guide::Net *net);
int guide_shade_view(wr_guide *g, int set, float alpha, float record_prob, bool collect, uint32_t max_bounce,
                     wr::guide::Shade *out);
// Clears the record set and picks the fraction of paths to record so that
// about 90% of the capacity fills.
int guide_begin_pass(wr_guide *g, int set, uint64_t samples, float *record_prob);
int guide_end_pass(wr_guide *g, int set, const float4 *sample_rad, uint64_t samples, float record_prob, uint32_t steps,
                   wr_guide_stats *gs);
int guide_render_end(wr_guide *g, wr_guide_stats *gs);
// Sizes the per-entry draws of the inference kernel for a path pool of `capacity` entries.
int guide_reserve_draws(wr_guide *g, uint32_t capacity);
wr_scene *guide_scene(const wr_guide *g);
