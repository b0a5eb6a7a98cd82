// wr_debug.hip -- per-call evaluation of the device BSDF, light and camera
// functions on host inputs.
//
// The renders exercise these functions only in aggregate; these entry points
// run the very functions the kernels inline (wr_device.h: bsdf_eval, bsdf_pdf,
// bsdf_sample, sample_light, lens_point_u, primary_dir_u, camera_we1) one
// record per lane, so each can be compared call by call with the reference's
// own classes (bsdf/*.h, light/diffusearealight.h:53-72, camera/thinlens.h:62-166).
// Diagnostics only: inputs are staged through device buffers allocated per call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/wurst_amd.h"
#include "wr_device.h"
#include "wr_host.h"

namespace {

constexpr int kDebugBlock = 256;

__device__ __forceinline__ float3 ld3(const float *p, uint32_t i) { return f3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
__device__ __forceinline__ void st3(float *p, uint32_t i, float3 v) {
    p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z;
}

// Bsdf::evalBsdf and evalPdfW (bsdf.h:20-42) in the local frame (y = normal).
__global__ void k_debug_bsdf_eval(const DMat *mats, const float *wi, const float *wo, uint32_t n, float *f, float *pdf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DMat m = mats[i];
    const float3 a = ld3(wi, i), b = ld3(wo, i);
    st3(f, i, bsdf_eval(m, a, b));
    pdf[i] = bsdf_pdf(m, a, b);
}

// Bsdf::sampleBsdfCosByPdf (bsdf.h:20-42): direction, f cos / pdf and pdf.
__global__ void k_debug_bsdf_sample(const DMat *mats, const float *wi, const float *u, uint32_t n, float *wo, float *w,
                                    float *pdf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DMat m = mats[i];
    float3 o = f3(0.f, 0.f, 0.f);
    float p = 0.f;
    const float3 v = bsdf_sample(m, ld3(wi, i), u[2 * i], u[2 * i + 1], o, p);
    st3(wo, i, o);
    st3(w, i, v);
    pdf[i] = p;
}

// createLightSubpath(N = 1): light choice, point, normal, contribution and pdfA.
__global__ void k_debug_light(DScene s, const float *u, uint32_t n, float *pos, float *nrm, float *contrib, float *pdfA) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const LightSample l = sample_light<false>(s, u[2 * i], u[2 * i + 1], 0u);
    st3(pos, i, l.pos);
    st3(nrm, i, l.n);
    st3(contrib, i, l.contrib);
    pdfA[i] = l.pdfA;
}

// sampleWe0ByPdf + samplePerPixelWe1CosByPdf: origin, direction and pdfW of a
// camera sample at pixel (px, py) with lens and jitter uniforms.
__global__ void k_debug_camera_ray(DCamera cam, const uint32_t *pixel, const float *ulens, const float *upix, uint32_t n,
                                   float *org, float *dir, float *pdfW) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float3 lens = lens_point_u(cam, ulens[2 * i], ulens[2 * i + 1]);
    float p = 0.f;
    const float3 d = primary_dir_u(cam, upix[2 * i], upix[2 * i + 1], pixel[2 * i], pixel[2 * i + 1], lens, p);
    st3(org, i, cam_to_world(cam, lens) + f3(cam.ox, cam.oy, cam.oz));
    st3(dir, i, d);
    pdfW[i] = p;
}

// ThinlensCamera::evalWe1 toward direction w from the lens point `org` (world).
__global__ void k_debug_camera_we1(DCamera cam, const float *org, const float *w, uint32_t n, float *we1, float *raster) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float3 lens = cam_to_local(cam, ld3(org, i) - f3(cam.ox, cam.oy, cam.oz));
    float rx = -1.f, ry = -1.f;
    we1[i] = camera_we1(cam, lens, ld3(w, i), rx, ry);
    raster[2 * i] = rx;
    raster[2 * i + 1] = ry;
}

// Device staging for one call: inputs copied in, outputs copied out, all freed.
struct Staging {
    std::vector<void *> ptrs;
    ~Staging() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int in(T **dst, const void *src, size_t bytes) {
        *dst = nullptr;
        WR_HIP(hipMalloc(reinterpret_cast<void **>(dst), bytes ? bytes : 4));
        ptrs.push_back(*dst);
        if (bytes) WR_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return WR_OK;
    }
    template <class T>
    int out(T **dst, size_t bytes) {
        *dst = nullptr;
        WR_HIP(hipMalloc(reinterpret_cast<void **>(dst), bytes ? bytes : 4));
        ptrs.push_back(*dst);
        return WR_OK;
    }
};

int to_dmats(const wr_material *m, uint32_t n, std::vector<DMat> &out) {
    out.resize(n);
    for (uint32_t i = 0; i < n; i++) {
        if (m[i].type < WR_LAMBERT || m[i].type > WR_MIXED) return set_error(WR_ERR_INVALID_ARGUMENT, "bad material type");
        DMat &d = out[i];
        d = DMat{};
        d.type = m[i].type;
        d.twosided = m[i].twosided;
        d.sigma = m[i].sigma;
        d.n = m[i].exponent;
        for (int k = 0; k < 3; k++) { d.rho_d[k] = m[i].rho_d[k]; d.rho_s[k] = m[i].rho_s[k]; }
    }
    return WR_OK;
}

inline unsigned blocks(uint32_t n) { return (n + kDebugBlock - 1) / kDebugBlock; }

// Scattered 16-byte loads: lane l's k-th load reads the first 16 bytes of line
// (h_k >> 8) & line_mask of a buffer of line_mask + 1 lines of 64 bytes (in
// bounds by the mask), eight independent loads per step; the xor of
// everything read decides an unlikely store, so no load can be dropped.
__global__ __launch_bounds__(kDebugBlock) void k_debug_l1_probe(const uint4 *__restrict__ buf, uint32_t line_mask,
                                                                uint32_t steps, uint32_t *sink) {
    uint32_t h = (blockIdx.x * kDebugBlock + threadIdx.x) * 0x9E3779B9u + 0x7F4A7C15u;
    uint32_t acc = 0;
    for (uint32_t i = 0; i < steps; i++) {
        uint4 v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            h = h * 1664525u + 1013904223u;
            v[k] = buf[size_t((h >> 8) & line_mask) * 4];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) acc ^= v[k].x ^ v[k].y ^ v[k].z ^ v[k].w;
    }
    if (acc == 0x9E3779B9u) sink[0] = acc;
}

}  // namespace

int wr_debug_bsdf_eval(wr_device *dev, const wr_material *mats, const float *wi, const float *wo, uint32_t n, float *f,
                       float *pdf) {
    if (!dev || (n && (!mats || !wi || !wo || !f || !pdf))) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0) return WR_OK;
    std::vector<DMat> dm;
    if (int rc = to_dmats(mats, n, dm)) return rc;
    if (int rc = set_device(dev)) return rc;
    Staging S;
    DMat *d_m; float *d_wi, *d_wo, *d_f, *d_p;
    const size_t v3 = sizeof(float) * 3 * size_t(n);
    if (int rc = S.in(&d_m, dm.data(), sizeof(DMat) * n)) return rc;
    if (int rc = S.in(&d_wi, wi, v3)) return rc;
    if (int rc = S.in(&d_wo, wo, v3)) return rc;
    if (int rc = S.out(&d_f, v3)) return rc;
    if (int rc = S.out(&d_p, sizeof(float) * n)) return rc;
    k_debug_bsdf_eval<<<blocks(n), kDebugBlock>>>(d_m, d_wi, d_wo, n, d_f, d_p);
    WR_HIP(hipGetLastError());
    WR_HIP(hipMemcpy(f, d_f, v3, hipMemcpyDeviceToHost));
    WR_HIP(hipMemcpy(pdf, d_p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return WR_OK;
}

int wr_debug_bsdf_sample(wr_device *dev, const wr_material *mats, const float *wi, const float *u, uint32_t n, float *wo,
                         float *weight, float *pdf) {
    if (!dev || (n && (!mats || !wi || !u || !wo || !weight || !pdf)))
        return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0) return WR_OK;
    std::vector<DMat> dm;
    if (int rc = to_dmats(mats, n, dm)) return rc;
    if (int rc = set_device(dev)) return rc;
    Staging S;
    DMat *d_m; float *d_wi, *d_u, *d_wo, *d_w, *d_p;
    const size_t v3 = sizeof(float) * 3 * size_t(n);
    if (int rc = S.in(&d_m, dm.data(), sizeof(DMat) * n)) return rc;
    if (int rc = S.in(&d_wi, wi, v3)) return rc;
    if (int rc = S.in(&d_u, u, sizeof(float) * 2 * size_t(n))) return rc;
    if (int rc = S.out(&d_wo, v3)) return rc;
    if (int rc = S.out(&d_w, v3)) return rc;
    if (int rc = S.out(&d_p, sizeof(float) * n)) return rc;
    k_debug_bsdf_sample<<<blocks(n), kDebugBlock>>>(d_m, d_wi, d_u, n, d_wo, d_w, d_p);
    WR_HIP(hipGetLastError());
    WR_HIP(hipMemcpy(wo, d_wo, v3, hipMemcpyDeviceToHost));
    WR_HIP(hipMemcpy(weight, d_w, v3, hipMemcpyDeviceToHost));
    WR_HIP(hipMemcpy(pdf, d_p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return WR_OK;
}

int wr_debug_light_sample(wr_scene *s, const float *u, uint32_t n, float *pos, float *nrm, float *contrib, float *pdfA) {
    if (!s || (n && (!u || !pos || !nrm || !contrib || !pdfA))) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    if (s->ds.n_lights == 0) return set_error(WR_ERR_INVALID_ARGUMENT, "scene has no area lights");
    if (n == 0) return WR_OK;
    if (int rc = set_device(s->dev)) return rc;
    Staging S;
    float *d_u, *d_pos, *d_n, *d_c, *d_p;
    const size_t v3 = sizeof(float) * 3 * size_t(n);
    if (int rc = S.in(&d_u, u, sizeof(float) * 2 * size_t(n))) return rc;
    if (int rc = S.out(&d_pos, v3)) return rc;
    if (int rc = S.out(&d_n, v3)) return rc;
    if (int rc = S.out(&d_c, v3)) return rc;
    if (int rc = S.out(&d_p, sizeof(float) * n)) return rc;
    WR_HIP(hipStreamSynchronize(s->stream));
    k_debug_light<<<blocks(n), kDebugBlock>>>(s->ds, d_u, n, d_pos, d_n, d_c, d_p);
    WR_HIP(hipGetLastError());
    WR_HIP(hipMemcpy(pos, d_pos, v3, hipMemcpyDeviceToHost));
    WR_HIP(hipMemcpy(nrm, d_n, v3, hipMemcpyDeviceToHost));
    WR_HIP(hipMemcpy(contrib, d_c, v3, hipMemcpyDeviceToHost));
    WR_HIP(hipMemcpy(pdfA, d_p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return WR_OK;
}

int wr_debug_camera_ray(wr_scene *s, const uint32_t *pixel, const float *ulens, const float *upix, uint32_t n, float *org,
                        float *dir, float *pdfW) {
    if (!s || (n && (!pixel || !ulens || !upix || !org || !dir || !pdfW)))
        return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    for (uint32_t i = 0; i < n; i++)
        if (pixel[2 * i] >= uint32_t(s->W) || pixel[2 * i + 1] >= uint32_t(s->H))
            return set_error(WR_ERR_INVALID_ARGUMENT, "pixel outside the film");
    if (n == 0) return WR_OK;
    if (int rc = set_device(s->dev)) return rc;
    Staging S;
    uint32_t *d_px;
    float *d_ul, *d_up, *d_o, *d_d, *d_p;
    const size_t v2 = sizeof(float) * 2 * size_t(n), v3 = sizeof(float) * 3 * size_t(n);
    if (int rc = S.in(&d_px, pixel, sizeof(uint32_t) * 2 * size_t(n))) return rc;
    if (int rc = S.in(&d_ul, ulens, v2)) return rc;
    if (int rc = S.in(&d_up, upix, v2)) return rc;
    if (int rc = S.out(&d_o, v3)) return rc;
    if (int rc = S.out(&d_d, v3)) return rc;
    if (int rc = S.out(&d_p, sizeof(float) * n)) return rc;
    k_debug_camera_ray<<<blocks(n), kDebugBlock>>>(s->ds.cam, d_px, d_ul, d_up, n, d_o, d_d, d_p);
    WR_HIP(hipGetLastError());
    WR_HIP(hipMemcpy(org, d_o, v3, hipMemcpyDeviceToHost));
    WR_HIP(hipMemcpy(dir, d_d, v3, hipMemcpyDeviceToHost));
    WR_HIP(hipMemcpy(pdfW, d_p, sizeof(float) * n, hipMemcpyDeviceToHost));
    return WR_OK;
}

int wr_debug_camera_we1(wr_scene *s, const float *org, const float *w, uint32_t n, float *we1, float *raster) {
    if (!s || (n && (!org || !w || !we1 || !raster))) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0) return WR_OK;
    if (int rc = set_device(s->dev)) return rc;
    Staging S;
    float *d_o, *d_w, *d_we, *d_r;
    const size_t v3 = sizeof(float) * 3 * size_t(n);
    if (int rc = S.in(&d_o, org, v3)) return rc;
    if (int rc = S.in(&d_w, w, v3)) return rc;
    if (int rc = S.out(&d_we, sizeof(float) * n)) return rc;
    if (int rc = S.out(&d_r, sizeof(float) * 2 * size_t(n))) return rc;
    k_debug_camera_we1<<<blocks(n), kDebugBlock>>>(s->ds.cam, d_o, d_w, n, d_we, d_r);
    WR_HIP(hipGetLastError());
    WR_HIP(hipMemcpy(we1, d_we, sizeof(float) * n, hipMemcpyDeviceToHost));
    WR_HIP(hipMemcpy(raster, d_r, sizeof(float) * 2 * size_t(n), hipMemcpyDeviceToHost));
    return WR_OK;
}

int wr_debug_l1_request_probe(wr_device *dev, uint64_t footprint, uint32_t loads_per_lane, double *ms, uint64_t *loads,
                              uint32_t *n_cu) {
    if (!dev || !ms || !loads || !n_cu) return set_error(WR_ERR_INVALID_ARGUMENT, "null argument");
    // the line index takes 24 random bits: 2^24 lines of 64 B
    if (footprint < 64 || footprint > (uint64_t(1) << 30) || (footprint & (footprint - 1)))
        return set_error(WR_ERR_INVALID_ARGUMENT, "footprint must be a power of two in 64 B .. 1 GiB");
    if (loads_per_lane < 8 || loads_per_lane % 8)
        return set_error(WR_ERR_INVALID_ARGUMENT, "loads_per_lane must be a positive multiple of 8");
    if (int rc = set_device(dev)) return rc;
    int per_cu = 0;
    WR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(&k_debug_l1_probe),
                                                        kDebugBlock, 0));
    const uint32_t grid = uint32_t(dev->n_cu) * uint32_t(std::max(1, per_cu));
    Staging S;
    uint4 *d_buf;
    uint32_t *d_sink;
    if (int rc = S.out(&d_buf, size_t(footprint))) return rc;
    if (int rc = S.out(&d_sink, sizeof(uint32_t))) return rc;
    WR_HIP(hipMemset(d_buf, 0, size_t(footprint)));
    const uint32_t mask = uint32_t(footprint / 64 - 1), steps = loads_per_lane / 8;
    hipEvent_t e0, e1;
    WR_HIP(hipEventCreate(&e0));
    WR_HIP(hipEventCreate(&e1));
    k_debug_l1_probe<<<grid, kDebugBlock>>>(d_buf, mask, steps, d_sink);  // warm-up (caches, clocks)
    (void)hipEventRecord(e0, nullptr);
    k_debug_l1_probe<<<grid, kDebugBlock>>>(d_buf, mask, steps, d_sink);
    (void)hipEventRecord(e1, nullptr);
    const hipError_t err = hipEventSynchronize(e1);
    float t = 0.f;
    if (err == hipSuccess) (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    WR_HIP(err);
    WR_HIP(hipGetLastError());
    *ms = double(t);
    *loads = uint64_t(grid) * kDebugBlock * loads_per_lane;
    *n_cu = uint32_t(dev->n_cu);
    return WR_OK;
}
