// wr_device.h -- device-side math, RNG, sampling and BSDFs (FP32).
//
// Each function names the reference code whose semantics it reproduces; the
// arithmetic is FP32 (the reference is FP64 throughout, README.md:83).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "wr_layout.h"

namespace wr {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kRayEps = 1e-4f;                // Ray3::Epsilon (src/common/math/ray.h:8)
constexpr float kOneMinusEps = 0x1.fffffep-1f;  // CdfTable remap clamp 1 - 1e-12 in FP32 (cdftable.h:41)

enum : int { kLambert = 0, kOrenNayar = 1, kPhong = 2, kMixed = 3 };

struct DMat {  // 48 B
    int type, twosided;
    float sigma, n;
    float rho_d[4];
    float rho_s[4];
};

struct DLight {  // 48 B
    float L[3];
    float pdf_select;   // CdfTable pdf of this light (weights = |L|_2, diffusearealight.h:21-24)
    float cdf_select;   // CdfTable cdf
    float inv_area;
    float area;
    uint32_t first_ltri;
    uint32_t n_ltri;
    uint32_t pad[3];
};

struct DCamera {
    float ox, oy, oz;
    float xx, xy, xz, yx, yy, yz, zx, zy, zz;  // frame (thinlens.h:46-50)
    float d;                                     // lens-to-film distance (camera.h:16)
    int W, H;
    float invW, invH;
    float per_pixel_norm, per_film_norm;
    float aperture, focal;
};

// Homogeneous medium filling the scene (homogeneous.h:11-66) with a
// Henyey-Greenstein phase function; on == 0: vacuum.
struct DMedium {
    float sig_t[3], sig_s[3];
    float g;
    int on;
};

struct DScene {
    const NodeQ *nodes;       // traversal nodes (quantised copy of the BVH4)
    const Node4 *fnodes;      // the BVH4 itself (per-pixel kernels)
    const float4 *tris;       // 3 float4 per leaf-order triangle
    const int2 *geom_info;    // (material, light) per geometry
    const DMat *mats;
    const DLight *lights;
    const float4 *ltris;      // 3 float4 per emitter triangle
    uint32_t n_lights;
    DCamera cam;
    DMedium med;
    // envmap light (EnvmapLight over an equirectangular image, light/envmap.h):
    // texels [env_h][env_w] (w = 2h), the texel CdfTable's cdf and pdf; env_select =
    // its light-selection pdf (0: no envmap)
    const float4 *env_tex;
    const float *env_cdf, *env_pdf;
    int env_w, env_h;
    float env_select;
};

// ---------------------------------------------------------------- float3 helpers
__device__ __forceinline__ float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 operator-(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 operator*(float3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 operator*(float3 a, float3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float3 operator-(float3 a) { return f3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float3 normalize3(float3 a) { return a * (1.0f / sqrtf(dot3(a, a))); }
__device__ __forceinline__ float3 xyz(float4 v) { return f3(v.x, v.y, v.z); }
__device__ __forceinline__ bool is_black(float3 c) { return c.x == 0.f && c.y == 0.f && c.z == 0.f; }

// ---------------------------------------------------------------- RNG
// Counter-based: a pure integer function of (seed, pixel, sample, dim); the
// oracle computes the same integers.  Dimension layout per path: 0-1 light,
// 2-3 lens, 4-5 pixel jitter, then 6 per bounce (0-1 medium, 2-3 BSDF).
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ uint32_t path_key(uint32_t seed, uint32_t pixel, uint32_t sample) {
    uint32_t h = mix32(seed ^ 0xa511e9b3u);
    h = mix32(h ^ pixel);
    return mix32(h ^ (sample * 0x9e3779b9u + 0x7f4a7c15u));
}
__device__ __forceinline__ float u01(uint32_t key, uint32_t dim) {
    return float(mix32(key + dim * 0x9e3779b9u + 0x632be5abu) >> 8) * 0x1p-24f;
}
enum : uint32_t { kDimLight = 0, kDimLens = 2, kDimPixel = 4, kDimBounce0 = 6, kDimsPerBounce = 6, kBdMedium = 0, kBdBsdf = 2 };
// position inside the envmap texel of a light sample: fresh numbers (re-using the
// twice-remapped selection number, as the FP64 reference does, would leave FP32
// a few bits of it)
constexpr uint32_t kDimEnv = 0x00fffff0u;

// ---------------------------------------------------------------- frames and warps
struct Frame { float3 x, y, z; };
// CoordFrame3(normal): Duff et al. 2017 ONB with y = normal (coordframe.h:46-60)
__device__ __forceinline__ Frame make_frame(float3 n) {
    Frame f;
    f.y = n;
    float sign = copysignf(1.0f, n.y);
    float a = -1.0f / (sign + n.y);
    float b = n.z * n.x * a;
    f.x = f3(sign + n.x * n.x * a, -n.x, b);
    f.z = f3(sign * b, -sign * n.z, 1.0f + sign * n.z * n.z * a);
    return f;
}
__device__ __forceinline__ float3 to_local(const Frame &f, float3 v) { return f3(dot3(v, f.x), dot3(v, f.y), dot3(v, f.z)); }
__device__ __forceinline__ float3 to_world(const Frame &f, float3 v) { return f.x * v.x + f.y * v.y + f.z * v.z; }
__device__ __forceinline__ float3 flipy(float3 v) { return f3(v.x, -v.y, v.z); }
// Mapping::CosineWeightedHemisphereFromSquare (mapping.h:42-49)
__device__ __forceinline__ float3 cosine_hemisphere(float u0, float u1) {
    float s, c;
    sincospif(2.0f * u1, &s, &c);
    float sinphi = sqrtf(1.0f - u0);
    return f3(c * sinphi, sqrtf(u0), s * sinphi);
}
// Mapping::DiskFromSquare, concentric map (mapping.h:53-73)
__device__ __forceinline__ void disk_from_square(float u0, float u1, float &dx, float &dy) {
    float a = u0 * 2.0f - 1.0f, b = u1 * 2.0f - 1.0f;
    if (a == 0.f && b == 0.f) { dx = 0.f; dy = 0.f; return; }
    float r, phi;
    if (a * a > b * b) { r = a; phi = 0.25f * (b / a); }
    else { r = b; phi = 0.5f - 0.25f * (a / b); }
    float s, c;
    sincospif(phi, &s, &c);
    dx = r * c; dy = r * s;
}

// ---------------------------------------------------------------- BSDFs (local frame, y = normal)
__device__ __forceinline__ float3 rho_d(const DMat &m) { return f3(m.rho_d[0], m.rho_d[1], m.rho_d[2]); }
__device__ __forceinline__ float3 rho_s(const DMat &m) { return f3(m.rho_s[0], m.rho_s[1], m.rho_s[2]); }
// LambertDiffuseBsdf (lambertdiffuse.h:16-51)
__device__ __forceinline__ float3 lambert_eval(const DMat &m, float3 wi, float3 wo) {
    if (wi.y <= 0.f || wo.y <= 0.f) return f3(0, 0, 0);
    return rho_d(m) * kInvPi;
}
__device__ __forceinline__ float lambert_pdf(float3 wi, float3 wo) {
    if (wi.y <= 0.f || wo.y <= 0.f) return 0.f;
    return wo.y * kInvPi;
}
// OrenNayarBsdf (orennayar.h:14-83)
__device__ __forceinline__ float3 oren_eval(const DMat &m, float3 wi, float3 wo) {
    if (wi.y <= 0.f || wo.y <= 0.f) return f3(0, 0, 0);
    float sinI = sqrtf(fminf(fmaxf(1.f - wi.y * wi.y, 0.f), 1.f));
    float sinO = sqrtf(fminf(fmaxf(1.f - wo.y * wo.y, 0.f), 1.f));
    float maxCos = 0.f, sinAlpha = 0.f, tanBeta = 0.f;
    if (sinI > 1e-4f && sinO > 1e-4f) {
        float dcos = (wi.x / sinI) * (wo.x / sinO) + (wi.z / sinI) * (wo.z / sinO);
        maxCos = fmaxf(0.f, dcos);
        if (fabsf(wi.y) > fabsf(wo.y)) { sinAlpha = sinO; tanBeta = sinI / wi.y; }
        else { sinAlpha = sinI; tanBeta = sinO / wo.y; }
    }
    float s2 = m.sigma * m.sigma;
    float A = 1.f - s2 / (2.f * (s2 + 0.33f));
    float B = 0.45f * s2 / (s2 + 0.09f);
    return rho_d(m) * (kInvPi * (A + B * maxCos * sinAlpha * tanBeta));
}
// PhongBsdf, modified Phong (phong.h:22-93)
__device__ __forceinline__ float3 phong_eval(const DMat &m, float3 wi, float3 wo) {
    if (wi.y <= 0.f || wo.y <= 0.f) return f3(0, 0, 0);
    float3 r = f3(-wi.x, wi.y, -wi.z);
    if (r.y <= 0.f) return f3(0, 0, 0);
    float c = dot3(r, wo);
    if (c <= 0.f) return f3(0, 0, 0);
    return rho_s(m) * ((m.n + 2.f) * powf(c, m.n) * (kInvPi * 0.5f));
}
__device__ __forceinline__ float phong_pdf(const DMat &m, float3 wi, float3 wo) {
    if (wi.y <= 0.f && wo.y <= 0.f) return 0.f;
    float3 r = f3(-wi.x, wi.y, -wi.z);
    float c = dot3(r, wo);
    if (c <= 0.f) return 0.f;
    return (m.n + 1.f) * powf(c, m.n) * (kInvPi * 0.5f);
}
__device__ __forceinline__ float3 phong_dir(const DMat &m, float3 wi, float u0, float u1) {
    float3 r = f3(-wi.x, wi.y, -wi.z);
    float cosT = powf(u0, 1.0f / (m.n + 1.0f));
    float sinT = sqrtf(fmaxf(1.0f - cosT * cosT, 0.f));
    float s, c;
    sincospif(2.0f * u1, &s, &c);
    Frame rb = make_frame(r);
    return to_world(rb, f3(sinT * c, cosT, sinT * s));
}
__device__ __forceinline__ float3 inner_eval(const DMat &m, float3 wi, float3 wo) {
    switch (m.type) {
    case kLambert: return lambert_eval(m, wi, wo);
    case kOrenNayar: return oren_eval(m, wi, wo);
    case kPhong: return phong_eval(m, wi, wo);
    default: return lambert_eval(m, wi, wo) + phong_eval(m, wi, wo);  // MixedBsdf (mixed.h:20-29)
    }
}
__device__ __forceinline__ float inner_pdf(const DMat &m, float3 wi, float3 wo) {
    switch (m.type) {
    case kLambert: case kOrenNayar: return lambert_pdf(wi, wo);
    case kPhong: return phong_pdf(m, wi, wo);
    default: return 0.5f * (lambert_pdf(wi, wo) + phong_pdf(m, wi, wo));  // mixed.h:31-40
    }
}
// sampleBsdfCosByPdf of each BSDF: returns f*cos/pdf, writes wo and pdf
__device__ __forceinline__ float3 inner_sample(const DMat &m, float3 wi, float u0, float u1, float3 &wo, float &pdf) {
    const float3 zero = f3(0, 0, 0);
    switch (m.type) {
    case kLambert: {
        if (wi.y <= 0.f) return zero;
        float3 o = cosine_hemisphere(u0, u1);
        wo = o;
        pdf = fmaxf(o.y, 0.f) * kInvPi;
        if (o.y <= 0.f) return zero;
        return rho_d(m);
    }
    case kOrenNayar: {
        if (wi.y < 0.f) { pdf = 0.f; return zero; }
        float3 o = cosine_hemisphere(u0, u1);
        wo = o;
        pdf = o.y * kInvPi;
        return oren_eval(m, o, wi) * kPi;
    }
    case kPhong: {
        if (wi.y <= 0.f) return zero;
        float3 r = f3(-wi.x, wi.y, -wi.z);
        float3 o = phong_dir(m, wi, u0, u1);
        float c = dot3(o, r);
        wo = o;
        pdf = (m.n + 1.f) * 0.5f * kInvPi * powf(c, m.n);
        if (r.y <= 0.f || o.y <= 0.f || c <= 0.f) return zero;
        return rho_s(m) * ((m.n + 2.f) / (m.n + 1.f) * o.y);
    }
    default: {  // MixedBsdf::sampleBsdfCosByPdf (mixed.h:42-60), two components
        float3 o = f3(0, 0, 0);
        float ci = fminf(floorf(u0 * 2.f), 1.f);
        float ur = fminf(u0 * 2.f - ci, kOneMinusEps);
        if (wi.y > 0.f) {
            if (ci == 0.f) o = cosine_hemisphere(ur, u1);
            else o = phong_dir(m, wi, ur, u1);
        }
        wo = o;
        float3 f = inner_eval(m, wi, o);
        float p = inner_pdf(m, wi, o);
        pdf = p;
        if (p == 0.f) return zero;
        return f * (o.y / p);
    }
    }
}
// TwoSidedBsdf wrapper (twosided.h:25-63)
__device__ __forceinline__ float3 bsdf_eval(const DMat &m, float3 wi, float3 wo) {
    if (!m.twosided) return inner_eval(m, wi, wo);
    if (wi.y > 0.f && wo.y > 0.f) return inner_eval(m, wi, wo);
    if (!(wi.y > 0.f) && !(wo.y > 0.f)) return inner_eval(m, flipy(wi), flipy(wo));
    return f3(0, 0, 0);
}
__device__ __forceinline__ float bsdf_pdf(const DMat &m, float3 wi, float3 wo) {
    if (!m.twosided) return inner_pdf(m, wi, wo);
    if (wi.y > 0.f && wo.y > 0.f) return inner_pdf(m, wi, wo);
    if (!(wi.y > 0.f) && !(wo.y > 0.f)) return inner_pdf(m, flipy(wi), flipy(wo));
    return 0.f;
}
__device__ __forceinline__ float3 bsdf_sample(const DMat &m, float3 wi, float u0, float u1, float3 &wo, float &pdf) {
    if (!m.twosided || wi.y > 0.f) return inner_sample(m, wi, u0, u1, wo, pdf);
    float3 r = inner_sample(m, flipy(wi), u0, u1, wo, pdf);
    wo = flipy(wo);
    return r;
}

// ---------------------------------------------------------------- lights
// env: an envmap vertex, whose pos is its direction (path.h:126-131)
struct LightSample { float3 pos, n, contrib; float pdfA; bool env; };

// ---------------------------------------------------------------- envmap (light/envmap.h, equirectangular.h)
// Mapping::PanoramaFromWorld / WorldFromPanorama (mapping.h:91-96,302-323)
__device__ __forceinline__ void panorama_from_world(float3 d, float &u, float &v) {
    u = atan2f(-d.z, -d.x) * (0.5f * kInvPi) + 0.5f;
    v = acosf(fminf(1.f, fmaxf(-1.f, d.y))) * kInvPi;
}
// Fimage::evalTexel, S repeat, T clamp (floatimage.h:91-112)
__device__ __forceinline__ float3 env_texel(const DScene &s, int x, int y) {
    x %= s.env_w;
    if (x < 0) x += s.env_w;
    y = min(max(y, 0), s.env_h - 1);
    return xyz(s.env_tex[size_t(y) * s.env_w + x]);
}
// EquiRectRep::eval = Fimage::evalBilinear (floatimage.h:120-133)
__device__ __forceinline__ float3 env_eval(const DScene &s, float3 d) {
    float u, v;
    panorama_from_world(d, u, v);
    const float sx = u * float(s.env_w) - 0.5f, sy = v * float(s.env_h) - 0.5f;
    const float fx = floorf(sx), fy = floorf(sy);
    const int x = int(fx), y = int(fy);
    const float dx1 = sx - fx, dy1 = sy - fy, dx2 = 1.f - dx1, dy2 = 1.f - dy1;
    return env_texel(s, x, y) * (dx2 * dy2) + env_texel(s, x, y + 1) * (dx2 * dy1) +
           env_texel(s, x + 1, y) * (dx1 * dy2) + env_texel(s, x + 1, y + 1) * (dx1 * dy1);
}
// EquiRectRep::pdfA: solid-angle pdf of sampling d (equirectangular.h:64-71)
__device__ __forceinline__ float env_pdf_dir(const DScene &s, float3 d) {
    float u, v;
    panorama_from_world(d, u, v);
    // clamped at both ends: a non-finite direction still reads inside the table
    const int x = min(max(int(u * float(s.env_w)), 0), s.env_w - 1), y = min(max(int(v * float(s.env_h)), 0), s.env_h - 1);
    const float sine = sqrtf(fmaxf(0.f, 1.f - d.y * d.y));
    if (!(sine > 1e-12f)) return 0.f;
    return s.env_pdf[size_t(y) * s.env_w + x] * float(s.env_w * s.env_h) / sine * (0.5f * kInvPi * kInvPi);
}
// EquiRectRep::sample (equirectangular.h:85-117): texel by the CdfTable
// (lower bound), (ox, oy) inside it; returns eval / pdf, or 0 where pdf = 0
__device__ __forceinline__ float3 env_sample(const DScene &s, float u, float ox, float oy, float3 &d, float &pdf) {
    uint32_t lo = 0, hi = uint32_t(s.env_w * s.env_h) - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s.env_cdf[mid] < u) lo = mid + 1; else hi = mid;
    }
    const int x = int(lo % uint32_t(s.env_w)), y = int(lo / uint32_t(s.env_w));
    const float th = (float(x) + ox) / float(s.env_w) * (2.f * kPi), ph = (float(y) + oy) / float(s.env_h) * kPi;
    float st, ct, sp, cp;
    sincosf(th, &st, &ct);
    sincosf(ph, &sp, &cp);
    d = f3(ct * sp, cp, st * sp);
    const float sine = sqrtf(fmaxf(0.f, 1.f - d.y * d.y));
    pdf = sine > 1e-12f ? s.env_pdf[lo] * float(s.env_w * s.env_h) / sine * (0.5f * kInvPi * kInvPi) : 0.f;
    return pdf > 0.f ? env_eval(s, d) * (1.f / pdf) : f3(0.f, 0.f, 0.f);
}
// createLightSubpath(numVertices = 1) -> CdfTable::sample -> DiffuseAreaLight::sampleLe0ByPdf
// -> TriangleMeshGeometry::sample (generalsubpathsampler.h:199-255, diffusearealight.h:53-72,
// trianglemesh.h:131-159, cdftable.h:36-44, mapping.h:21-27)
// ENV: the kernel variant for scenes with an envmap (the branch costs registers
// that the kernels of scenes without one keep).
template <bool ENV = false>
__device__ __forceinline__ LightSample sample_light(const DScene &s, float u0, float u1, uint32_t key) {
    LightSample ls;
    ls.env = false;
    if (ENV && s.env_select > 0.f && (s.n_lights == 0 || u0 > s.lights[s.n_lights - 1].cdf_select)) {
        // the envmap, last entry of the selection CdfTable (EnvmapLight::sampleLe0ByPdf, envmap.h:61-77)
        const float prev = s.n_lights == 0 ? 0.f : s.lights[s.n_lights - 1].cdf_select;
        const float r = fminf((u0 - prev) / s.env_select, kOneMinusEps);
        float pdfW;
        const float3 v = env_sample(s, r, u01(key, kDimEnv), u01(key, kDimEnv + 1), ls.pos, pdfW);
        ls.n = f3(0.f, 0.f, 0.f);
        ls.contrib = v * (1.f / s.env_select);
        ls.pdfA = s.env_select * pdfW;
        ls.env = true;
        return ls;
    }
    if (s.n_lights == 0) {  // nothing to sample: a zero contribution
        ls.pos = f3(0.f, 0.f, 0.f);
        ls.n = f3(0.f, 1.f, 0.f);
        ls.contrib = f3(0.f, 0.f, 0.f);
        ls.pdfA = 1.f;
        return ls;
    }
    uint32_t li = 0;
    while (li + 1 < s.n_lights && s.lights[li].cdf_select < u0) li++;
    const DLight &L = s.lights[li];
    float pdfL = L.pdf_select;
    float prev_cdf = li == 0 ? 0.f : s.lights[li - 1].cdf_select;
    float r = fminf((li == 0 ? u0 : (u0 - prev_cdf)) / pdfL, kOneMinusEps);
    // triangle by area: lower_bound over the per-light cdf stored in p0.w
    uint32_t lo = 0, hi = L.n_ltri - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (s.ltris[3 * (L.first_ltri + mid)].w < r) lo = mid + 1; else hi = mid;
    }
    const float4 p0 = s.ltris[3 * (L.first_ltri + lo) + 0];
    const float4 p1 = s.ltris[3 * (L.first_ltri + lo) + 1];
    const float4 p2 = s.ltris[3 * (L.first_ltri + lo) + 2];
    float prev_t = lo == 0 ? 0.f : s.ltris[3 * (L.first_ltri + lo - 1)].w;
    float r2 = fminf((lo == 0 ? r : (r - prev_t)) / p1.w, kOneMinusEps);
    float sq = sqrtf(r2);
    float b1 = sq * (1.f - u1), b2 = u1 * sq;
    float3 a = xyz(p0), b = xyz(p1), c = xyz(p2);
    ls.pos = a * (1.f - b1 - b2) + b * b1 + c * b2;
    ls.n = normalize3(cross3(b - a, c - a));
    float k = kPi * L.area / pdfL;  // (pi L / pdfPos) / pdfLight
    ls.contrib = f3(L.L[0] * k, L.L[1] * k, L.L[2] * k);
    ls.pdfA = pdfL * L.inv_area;
    return ls;
}

// ---------------------------------------------------------------- medium (homogeneous.h:11-66)
__device__ __forceinline__ float3 transmittance(const DMedium &m, float d) {
    return f3(expf(-m.sig_t[0] * d), expf(-m.sig_t[1] * d), expf(-m.sig_t[2] * d));
}
// Channel-averaged free-flight densities (homogeneous.h:36-50): passing a
// segment of length d, and ending the flight at distance d.
__device__ __forceinline__ float pass_pdf(const DMedium &m, float d) {
    const float3 t = transmittance(m, d);
    return (t.x + t.y + t.z) * (1.f / 3.f);
}
__device__ __forceinline__ float scatter_pdf(const DMedium &m, float d) {
    const float3 t = transmittance(m, d);
    return (m.sig_t[0] * t.x + m.sig_t[1] * t.y + m.sig_t[2] * t.z) * (1.f / 3.f);
}
// Free flight along a segment whose surface (or infinity) is t_surf away: the
// bounce's two medium draws pick a channel and a distance.  Returns true and
// the distance when the flight ends in the medium.
__device__ __forceinline__ bool free_flight(const DMedium &m, uint32_t key, uint32_t bounce, float t_surf, float &t) {
    const uint32_t dm = kDimBounce0 + kDimsPerBounce * bounce + kBdMedium;
    const float um = u01(key, dm), uc = u01(key, dm + 1);
    const int ch = min(int(uc * 3.f), 2);
    const float st = ch == 0 ? m.sig_t[0] : (ch == 1 ? m.sig_t[1] : m.sig_t[2]);
    const float tf = st > 0.f ? -logf(1.f - um) / st : INFINITY;
    t = tf;
    return tf < t_surf;
}
// Henyey-Greenstein (1941): cos between the propagation direction and the
// scattered direction; g > 0 scatters forward.
__device__ __forceinline__ float hg_pdf(float g, float c) {
    const float d = 1.f + g * g - 2.f * g * c;
    return (1.f - g * g) / (4.f * kPi * d * sqrtf(d));
}
// inversion of the HG CDF: cos = (1 + g^2 - s^2) / (2 g), s = (1 - g^2) / (1 - g + 2 g u)
__device__ __forceinline__ float3 hg_sample(float g, float3 fwd, float u0, float u1) {
    float c;
    if (fabsf(g) < 1e-3f) c = 1.f - 2.f * u0;
    else {
        const float q = (1.f - g * g) / (1.f - g + 2.f * g * u0);
        c = (1.f + g * g - q * q) / (2.f * g);
    }
    c = fminf(fmaxf(c, -1.f), 1.f);
    const float st = sqrtf(fmaxf(0.f, 1.f - c * c));
    float s, co;
    sincospif(2.f * u1, &s, &co);
    return normalize3(to_world(make_frame(fwd), f3(st * co, c, st * s)));
}

// ---------------------------------------------------------------- camera (thinlens.h, camera.h)
__device__ __forceinline__ float3 cam_to_world(const DCamera &c, float3 l) {
    return f3(c.xx * l.x + c.yx * l.y + c.zx * l.z, c.xy * l.x + c.yy * l.y + c.zy * l.z,
              c.xz * l.x + c.yz * l.y + c.zz * l.z);
}
__device__ __forceinline__ float3 cam_to_local(const DCamera &c, float3 w) {
    return f3(c.xx * w.x + c.xy * w.y + c.xz * w.z, c.yx * w.x + c.yy * w.y + c.yz * w.z,
              c.zx * w.x + c.zy * w.y + c.zz * w.z);
}
// ThinlensCamera::evalWe1 (thinlens.h:62-80): returns 0 outside the film
__device__ __forceinline__ float camera_we1(const DCamera &c, float3 lens_local, float3 w, float &rx, float &ry) {
    float3 dir = cam_to_local(c, w);
    float cz = dir.z;
    if (cz < 0.f) return 0.f;
    float3 fp = dir * (c.focal / cz) + lens_local;
    float3 film = fp * (-c.d / c.focal);
    rx = (film.x * c.invW + 0.5f) * float(c.W);
    ry = (film.y * c.invH + 0.5f) * float(c.H);
    if (!(rx >= 0.f && rx < float(c.W) && ry >= 0.f && ry < float(c.H))) return 0.f;
    float c2 = cz * cz;
    return c.per_film_norm / (c2 * c2);
}

// Lens point of a camera sample (sampleWe0ByPdf, thinlens.h:121-139): the
// pinhole, or the point of the aperture disk at uniforms (u0, u1).
__device__ __forceinline__ float3 lens_point_u(const DCamera &cam, float u0, float u1) {
    float lx = 0.f, ly = 0.f;
    if (cam.aperture > 0.f) {
        disk_from_square(u0, u1, lx, ly);
        lx *= cam.aperture; ly *= cam.aperture;
    }
    return f3(lx, ly, 0.f);
}
// Primary direction through pixel (px, py) jittered by (j0, j1)
// (samplePerPixelWe1CosByPdf, thinlens.h:141-166) in world space, and its
// per-pixel pdfW.
__device__ __forceinline__ float3 primary_dir_u(const DCamera &cam, float j0, float j1, uint32_t px, uint32_t py,
                                                float3 lens, float &pdfW) {
    const float ns0 = (float(px) + j0) * cam.invW, ns1 = (float(py) + j1) * cam.invH;
    const float3 film = f3((ns0 - 0.5f) * float(cam.W), (ns1 - 0.5f) * float(cam.H), -cam.d);
    const float3 focal = film * (cam.focal / film.z);
    const float3 dl = normalize3(focal - lens);
    pdfW = cam.per_pixel_norm / (dl.z * dl.z * dl.z);
    return cam_to_world(cam, dl);
}

}  // namespace wr
