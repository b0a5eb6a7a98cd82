/*
 * wr_oracle.c -- CPU FP64 restatement of the `nextevent_path_tracer` hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product.  See wr_oracle.h for what pins each function.
 *
 * Reference files restated here (paths relative to the reference checkout):
 *   src/renderer/pathnextevent.h          evalContribution / render
 *   src/pathsampler/generalsubpathsampler.h  randomWalk / createLightSubpath /
 *                                          createCameraSubpath / evalLightPdfA
 *   src/common/path.h                     GeometryTerm / ConvertToAreaFromSolidAngle
 *   src/common/scene.h                    transmittance (visibility part)
 *   src/accel/embreeaccel.h               FP32 ray casts, hit-vertex construction
 *   src/bsdf/{lambertdiffuse,orennayar,phong,mixed,twosided}.h
 *   src/light/diffusearealight.h, src/geometry/trianglemesh.h, src/common/cdftable.h
 *   src/camera/thinlens.h, src/common/camera.h, src/common/film.h
 *   src/common/mapping.h, src/common/math/{coordframe,local,ray}.h
 */
#define _GNU_SOURCE
#include "wr_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#define PI 3.141592653589793238462643383279502884197169399375105820974
#define INV_PI 0.318309886183790671537767526745028724068919291480912897495
#define SMALL_VALUE 1e-12 /* scalarmath.h:32 */
#define RAY_EPS 1e-4      /* ray.h:8 */

static __thread char g_err[256];
const char *orc_last_error(void) { return g_err; }
static int fail(const char *msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return 1;
}

/* ------------------------------------------------------------------ vec3 */
typedef struct { double x, y, z; } v3;
static inline v3 V(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, double s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vneg(v3 a) { return V(-a.x, -a.y, -a.z); }
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double vlen2(v3 a) { return vdot(a, a); }
static inline v3 vnorm(v3 a) { return vmul(a, 1.0 / sqrt(vlen2(a))); }
static inline double vget(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline v3 flipy(v3 a) { return V(a.x, -a.y, a.z); }

/* ------------------------------------------------------------------ RNG
 * Counter-based: uniform(seed, pixel, sample, dim).  A pure integer function,
 * so the GPU path draws bit-identical numbers.  The reference uses
 * mt19937_64 per tile (src/common/rng.h:5-29); the stream itself is not part
 * of the estimator, only the dimension order (one 2-D draw per decision). */
static inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
static inline uint32_t path_key(uint32_t seed, uint32_t pixel, uint32_t sample) {
    uint32_t h = mix32(seed ^ 0xa511e9b3u);
    h = mix32(h ^ pixel);
    h = mix32(h ^ (sample * 0x9e3779b9u + 0x7f4a7c15u));
    return h;
}
static inline double u01(uint32_t key, uint32_t dim) {
    return (double)(mix32(key + dim * 0x9e3779b9u + 0x632be5abu) >> 8) * (1.0 / 16777216.0);
}
uint32_t orc_rng_u32(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim) {
    return mix32(path_key(seed, pixel, sample) + dim * 0x9e3779b9u + 0x632be5abu);
}
double orc_rng_uniform(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim) {
    return u01(path_key(seed, pixel, sample), dim);
}
/* dimension layout per path */
enum { DIM_LIGHT = 0, DIM_LENS = 2, DIM_PIXEL = 4, DIM_BOUNCE0 = 6, DIMS_PER_BOUNCE = 6 };
/* position inside the envmap texel picked for a light sample (the reference
 * re-uses the twice-remapped selection number there; fresh numbers keep the
 * FP32 device's in-texel position at full resolution) */
enum { DIM_ENV = 0x00fffff0 };
enum { BD_MEDIUM = 0, BD_BSDF = 2 };

/* ------------------------------------------------------------------ frames, mappings */
typedef struct { v3 x, y, z; } frame3;
/* coordframe.h:46-60: Duff et al. 2017 ONB with y = normal */
static frame3 make_frame(v3 n) {
    frame3 f;
    f.y = n;
    double sign = copysign(1.0, n.y);
    double a = -1.0 / (sign + n.y);
    double b = n.z * n.x * a;
    f.x = V(sign + n.x * n.x * a, -n.x, b);
    f.z = V(sign * b, -sign * n.z, 1.0 + sign * n.z * n.z * a);
    return f;
}
static inline v3 to_local(const frame3 *f, v3 g) { return V(vdot(g, f->x), vdot(g, f->y), vdot(g, f->z)); }
static inline v3 to_world(const frame3 *f, v3 l) {
    return vadd(vadd(vmul(f->x, l.x), vmul(f->y, l.y)), vmul(f->z, l.z));
}
void orc_onb(const double n[3], double x[3], double z[3]) {
    frame3 f = make_frame(V(n[0], n[1], n[2]));
    x[0] = f.x.x; x[1] = f.x.y; x[2] = f.x.z;
    z[0] = f.z.x; z[1] = f.z.y; z[2] = f.z.z;
}
/* mapping.h:42-49 */
static v3 cosine_hemisphere(double u0, double u1) {
    double sinphi = sqrt(1.0 - u0);
    double theta = PI * 2.0 * u1;
    return V(cos(theta) * sinphi, sqrt(u0), sin(theta) * sinphi);
}
void orc_cosine_hemisphere(const double u[2], double out[3]) {
    v3 o = cosine_hemisphere(u[0], u[1]);
    out[0] = o.x; out[1] = o.y; out[2] = o.z;
}
/* mapping.h:21-27 */
static void barycentric_from_square(double u0, double u1, double *b1, double *b2) {
    double s = sqrt(u0);
    *b1 = s * (1.0 - u1);
    *b2 = u1 * s;
}
void orc_barycentric_from_square(const double u[2], double b[2]) { barycentric_from_square(u[0], u[1], &b[0], &b[1]); }
/* mapping.h:53-73 (concentric map) */
static void disk_from_square(double u0, double u1, double *dx, double *dy) {
    double a = u0 * 2.0 - 1.0, b = u1 * 2.0 - 1.0;
    if (a == 0 && b == 0) { *dx = 0; *dy = 0; return; }
    double r, phi;
    if (a * a > b * b) { r = a; phi = (PI / 4.0) * (b / a); }
    else { r = b; phi = (PI / 2.0) - (PI / 4.0) * (a / b); }
    *dx = r * cos(phi);
    *dy = r * sin(phi);
}
void orc_disk_from_square(const double u[2], double d[2]) { disk_from_square(u[0], u[1], &d[0], &d[1]); }

/* ------------------------------------------------------------------ CDF table (cdftable.h:10-44) */
typedef struct { int n; double *pdf, *cdf; } cdf_table;
static int cdf_build(cdf_table *t, const double *w, int n) {
    t->n = n; t->pdf = NULL; t->cdf = NULL;
    if (n == 0) return 0;
    t->pdf = (double *)malloc(sizeof(double) * n);
    t->cdf = (double *)malloc(sizeof(double) * n);
    if (!t->pdf || !t->cdf) return 1;
    double sum = 0.0;
    for (int i = 0; i < n; i++) sum += fabs(w[i]);
    for (int i = 0; i < n; i++) t->pdf[i] = fabs(w[i]) / sum;
    t->cdf[0] = fabs(w[0]) / sum;
    t->cdf[n - 1] = 1.0;
    for (int i = 1; i < n - 1; i++) t->cdf[i] = t->cdf[i - 1] + t->pdf[i];
    return 0;
}
static int cdf_sample(const cdf_table *t, double u, double *pdf, double *remapped) {
    /* std::lower_bound: first index whose cdf >= u */
    int lo = 0, hi = t->n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (t->cdf[mid] < u) lo = mid + 1; else hi = mid;
    }
    int idx = lo;
    if (idx >= t->n) idx = t->n - 1; /* unreachable for u < 1 (cdf[n-1] == 1) */
    if (pdf) *pdf = t->pdf[idx];
    if (remapped) {
        double r = (idx == 0) ? u / t->pdf[idx] : (u - t->cdf[idx - 1]) / t->pdf[idx];
        *remapped = fmin(r, 1.0 - SMALL_VALUE);
    }
    return idx;
}
static void cdf_free(cdf_table *t) { free(t->pdf); free(t->cdf); }
int orc_cdf_sample(const double *w, int n, double u, double *pdf, double *remapped) {
    cdf_table t;
    if (n <= 0 || cdf_build(&t, w, n)) return -1;
    int i = cdf_sample(&t, u, pdf, remapped);
    cdf_free(&t);
    return i;
}

/* ------------------------------------------------------------------ BSDFs (local frame, y = normal) */
typedef struct { double rho_d[3], rho_s[3], sigma, n; int type, twosided; } bsdf_t;
static void mat_load(bsdf_t *b, const orc_material *m) {
    for (int i = 0; i < 3; i++) { b->rho_d[i] = m->rho_d[i]; b->rho_s[i] = m->rho_s[i]; }
    b->sigma = m->sigma; b->n = m->exponent; b->type = m->type; b->twosided = m->twosided;
}
typedef struct { double r, g, b; } rgb;
static inline rgb RGB(double r, double g, double b) { rgb c = {r, g, b}; return c; }
static inline rgb rgb3(const double *a) { return RGB(a[0], a[1], a[2]); }
static inline rgb cmul(rgb a, rgb b) { return RGB(a.r * b.r, a.g * b.g, a.b * b.b); }
static inline rgb cscale(rgb a, double s) { return RGB(a.r * s, a.g * s, a.b * s); }
static inline rgb cadd(rgb a, rgb b) { return RGB(a.r + b.r, a.g + b.g, a.b + b.b); }
static inline int czero(rgb a) { return a.r == 0.0 && a.g == 0.0 && a.b == 0.0; } /* vecsimd.h:258-262 */
static const rgb BLACK = {0, 0, 0};

/* lambertdiffuse.h:16-51 */
static rgb lambert_eval(const bsdf_t *b, v3 wi, v3 wo) {
    if (wi.y <= 0.0 || wo.y <= 0.0) return BLACK;
    return cscale(rgb3(b->rho_d), INV_PI);
}
static double lambert_pdf(v3 wi, v3 wo) {
    if (wi.y <= 0.0 || wo.y <= 0.0) return 0.0;
    return wo.y * INV_PI;
}
static rgb lambert_sample(const bsdf_t *b, v3 wi, double u0, double u1, v3 *wo, double *pdf) {
    if (wi.y <= 0.0) return BLACK;
    v3 o = cosine_hemisphere(u0, u1);
    *wo = o;
    if (pdf) *pdf = fmax(o.y, 0.0) * INV_PI;
    if (o.y <= 0.0) return BLACK;
    return rgb3(b->rho_d);
}
/* orennayar.h:14-83 (pbrt-v3) */
static rgb oren_eval(const bsdf_t *b, v3 wi, v3 wo) {
    if (wi.y <= 0.0 || wo.y <= 0.0) return BLACK;
    double sinI = sqrt(fmin(fmax(1.0 - wi.y * wi.y, 0.0), 1.0));
    double sinO = sqrt(fmin(fmax(1.0 - wo.y * wo.y, 0.0), 1.0));
    double maxCos = 0.0, sinAlpha = 0.0, tanBeta = 0.0;
    if (sinI > 1e-4 && sinO > 1e-4) {
        double sinPI = wi.z / sinI, cosPI = wi.x / sinI;
        double sinPO = wo.z / sinO, cosPO = wo.x / sinO;
        double dcos = cosPI * cosPO + sinPI * sinPO;
        maxCos = fmax(0.0, dcos);
        if (fabs(wi.y) > fabs(wo.y)) { sinAlpha = sinO; tanBeta = sinI / wi.y; }
        else { sinAlpha = sinI; tanBeta = sinO / wo.y; }
    }
    double s2 = b->sigma * b->sigma;
    double A = 1.0 - s2 / (2 * (s2 + 0.33));
    double B = 0.45 * s2 / (s2 + 0.09);
    return cscale(rgb3(b->rho_d), INV_PI * (A + B * maxCos * sinAlpha * tanBeta));
}
static rgb oren_sample(const bsdf_t *b, v3 wi, double u0, double u1, v3 *wo, double *pdf) {
    if (wi.y < 0.0) { if (pdf) *pdf = 0.0; return BLACK; }
    v3 o = cosine_hemisphere(u0, u1);
    *wo = o;
    if (pdf) *pdf = o.y * INV_PI;
    return cscale(oren_eval(b, o, wi), PI); /* orennayar.h:74 passes (out, incoming) */
}
/* phong.h:22-93 (modified Phong, Lafortune & Willems 1994) */
static inline v3 reflect_local(v3 wi) { return V(-wi.x, wi.y, -wi.z); } /* Local3::Reflect(-wi) */
static rgb phong_eval(const bsdf_t *b, v3 wi, v3 wo) {
    if (wi.y <= 0.0 || wo.y <= 0.0) return BLACK;
    v3 r = reflect_local(wi);
    if (r.y <= 0.0) return BLACK;
    double c = vdot(r, wo);
    if (c <= 0.0) return BLACK;
    return cscale(rgb3(b->rho_s), (b->n + 2.0) * pow(c, b->n) * INV_PI * 0.5);
}
static double phong_pdf(const bsdf_t *b, v3 wi, v3 wo) {
    if (wi.y <= 0.0 && wo.y <= 0.0) return 0.0;
    v3 r = reflect_local(wi);
    double c = vdot(r, wo);
    if (c <= 0.0) return 0.0;
    return (b->n + 1.0) * pow(c, b->n) * INV_PI * 0.5;
}
static rgb phong_sample(const bsdf_t *b, v3 wi, double u0, double u1, v3 *wo, double *pdf) {
    if (wi.y <= 0.0) return BLACK;
    double n = b->n;
    v3 r = reflect_local(wi);
    double cosT = pow(u0, 1.0 / (n + 1.0));
    double sinT = sqrt(1.0 - cosT * cosT);
    double phi = 2.0 * PI * u1;
    v3 lo = V(sinT * cos(phi), cosT, sinT * sin(phi));
    frame3 rb = make_frame(r);
    v3 o = to_world(&rb, lo);
    double c = vdot(o, r);
    *wo = o;
    if (pdf) *pdf = (n + 1.0) * 0.5 * INV_PI * pow(c, n);
    if (r.y <= 0.0 || o.y <= 0.0 || c <= 0.0) return BLACK;
    return cscale(rgb3(b->rho_s), (n + 2.0) / (n + 1.0) * o.y);
}
/* one-sided dispatch (the material's own BSDF) */
static rgb inner_eval(const bsdf_t *b, v3 wi, v3 wo) {
    switch (b->type) {
    case ORC_LAMBERT: return lambert_eval(b, wi, wo);
    case ORC_OREN_NAYAR: return oren_eval(b, wi, wo);
    case ORC_PHONG: return phong_eval(b, wi, wo);
    default: return cadd(lambert_eval(b, wi, wo), phong_eval(b, wi, wo)); /* mixed.h:20-29 */
    }
}
static double inner_pdf(const bsdf_t *b, v3 wi, v3 wo) {
    switch (b->type) {
    case ORC_LAMBERT: case ORC_OREN_NAYAR: return lambert_pdf(wi, wo); /* orennayar.h:57-64 same form */
    case ORC_PHONG: return phong_pdf(b, wi, wo);
    default: return (lambert_pdf(wi, wo) + phong_pdf(b, wi, wo)) / 2.0; /* mixed.h:31-40 */
    }
}
static rgb inner_sample(const bsdf_t *b, v3 wi, double u0, double u1, v3 *wo, double *pdf) {
    switch (b->type) {
    case ORC_LAMBERT: return lambert_sample(b, wi, u0, u1, wo, pdf);
    case ORC_OREN_NAYAR: return oren_sample(b, wi, u0, u1, wo, pdf);
    case ORC_PHONG: return phong_sample(b, wi, u0, u1, wo, pdf);
    default: { /* mixed.h:42-60: two components, chosen uniformly */
        unsigned nc = 2;
        unsigned ci = (unsigned)floor(u0 * nc);
        if (ci > nc - 1) ci = nc - 1;
        double ur = fmin(u0 * nc - ci, 1.0 - SMALL_VALUE);
        v3 o = V(0, 0, 0);
        if (ci == 0) lambert_sample(b, wi, ur, u1, &o, NULL);
        else phong_sample(b, wi, ur, u1, &o, NULL);
        *wo = o;
        rgb f = inner_eval(b, wi, o);
        double p = inner_pdf(b, wi, o);
        if (pdf) *pdf = p;
        if (p == 0.0) return BLACK;
        return cscale(f, o.y / p);
    }
    }
}
/* twosided.h:25-63 wrapper */
static rgb bsdf_eval(const bsdf_t *b, v3 wi, v3 wo) {
    if (!b->twosided) return inner_eval(b, wi, wo);
    if (wi.y > 0.0 && wo.y > 0.0) return inner_eval(b, wi, wo);
    if (!(wi.y > 0.0) && !(wo.y > 0.0)) return inner_eval(b, flipy(wi), flipy(wo));
    return BLACK;
}
static double bsdf_pdf(const bsdf_t *b, v3 wi, v3 wo) {
    if (!b->twosided) return inner_pdf(b, wi, wo);
    if (wi.y > 0.0 && wo.y > 0.0) return inner_pdf(b, wi, wo);
    if (!(wi.y > 0.0) && !(wo.y > 0.0)) return inner_pdf(b, flipy(wi), flipy(wo));
    return 0.0;
}
static rgb bsdf_sample(const bsdf_t *b, v3 wi, double u0, double u1, v3 *wo, double *pdf) {
    if (!b->twosided || wi.y > 0.0) return inner_sample(b, wi, u0, u1, wo, pdf);
    rgb r = inner_sample(b, flipy(wi), u0, u1, wo, pdf);
    *wo = flipy(*wo);
    return r;
}
void orc_bsdf_eval(const orc_material *m, const double wi[3], const double wo[3], double f[3], double *pdf) {
    bsdf_t b; mat_load(&b, m);
    v3 i = V(wi[0], wi[1], wi[2]), o = V(wo[0], wo[1], wo[2]);
    rgb c = bsdf_eval(&b, i, o);
    f[0] = c.r; f[1] = c.g; f[2] = c.b;
    *pdf = bsdf_pdf(&b, i, o);
}
void orc_bsdf_eval_n(const orc_material *m, int n, const double *wi, const double *wo, double *f, double *pdf) {
    for (int i = 0; i < n; i++) orc_bsdf_eval(m, wi + 3 * i, wo + 3 * i, f + 3 * i, pdf + i);
}
void orc_bsdf_sample(const orc_material *m, const double wi[3], const double u[2], double wo[3], double w[3], double *pdf) {
    bsdf_t b; mat_load(&b, m);
    v3 o = V(0, 0, 0);
    double p = 0.0;
    rgb c = bsdf_sample(&b, V(wi[0], wi[1], wi[2]), u[0], u[1], &o, &p);
    wo[0] = o.x; wo[1] = o.y; wo[2] = o.z;
    w[0] = c.r; w[1] = c.g; w[2] = c.b;
    *pdf = p;
}

/* ------------------------------------------------------------------ scene + BVH */
typedef struct { double lo[3], hi[3]; int32_t left_or_first, count; } bnode; /* count>0: leaf */

typedef struct {
    int geom;
    rgb L;
    double area, inv_area;
    uint32_t first_tri, n_tris;
    cdf_table tri_cdf;
} light_t;

struct orc_scene {
    uint32_t n_verts, n_tris, n_geoms;
    v3 *verts;              /* FP32 vertices as doubles */
    uint32_t *tris;         /* [n_tris*3] */
    uint32_t *tri_geom;
    uint32_t *geom_first;   /* first triangle of each geometry */
    int32_t *geom_material, *geom_light;
    bsdf_t *mats; uint32_t n_mats;
    light_t *lights; uint32_t n_lights;
    cdf_table light_cdf;
    /* BVH */
    bnode *nodes; uint32_t n_nodes;
    uint32_t *tri_order;    /* leaf order -> triangle index */
    /* camera (thinlens.h:40-55, camera.h:9-17) */
    v3 cam_origin; frame3 cam_frame;
    int W, H;
    double cam_d, per_film_norm, per_pixel_norm, lens_area, aperture, focal;
    /* homogeneous medium filling the scene (homogeneous.h:11-66), HG phase */
    int has_medium;
    double sig_t[3], sig_s[3], hg_g;
    /* envmap (EnvmapLight over EquiRectRep, light/envmap.h, equirectangular.h):
     * texels [env_h][env_w] RGB, pixel CDF; last entry of light_cdf */
    int env_on, env_w, env_h;
    double *env_rgb;
    cdf_table env_cdf;
    double *light_w;  /* area-light CDF weights (|L|_2), for rebuilding light_cdf */
};

typedef struct { double c[3]; uint32_t tri; } bref;

static void box_tri(const orc_scene *s, uint32_t t, double lo[3], double hi[3]) {
    for (int k = 0; k < 3; k++) { lo[k] = DBL_MAX; hi[k] = -DBL_MAX; }
    for (int j = 0; j < 3; j++) {
        v3 p = s->verts[s->tris[3 * t + j]];
        for (int k = 0; k < 3; k++) {
            double v = vget(p, k);
            if (v < lo[k]) lo[k] = v;
            if (v > hi[k]) hi[k] = v;
        }
    }
}
static double half_area(const double lo[3], const double hi[3]) {
    double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0;
    return dx * dy + dy * dz + dz * dx;
}

typedef struct { bnode *nodes; uint32_t n, cap; } nodevec;
static uint32_t push_node(nodevec *nv) {
    if (nv->n == nv->cap) {
        nv->cap = nv->cap ? nv->cap * 2 : 1024;
        nv->nodes = (bnode *)realloc(nv->nodes, sizeof(bnode) * nv->cap);
    }
    return nv->n++;
}

#define NBINS 16
/* Binned SAH build (test-side BVH; any correct BVH serves the oracle). */
static void build_rec(const orc_scene *s, bref *refs, double (*tlo)[3], double (*thi)[3],
                      uint32_t begin, uint32_t end, nodevec *nv, uint32_t node) {
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    double clo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, chi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (uint32_t i = begin; i < end; i++) {
        uint32_t t = refs[i].tri;
        for (int k = 0; k < 3; k++) {
            if (tlo[t][k] < lo[k]) lo[k] = tlo[t][k];
            if (thi[t][k] > hi[k]) hi[k] = thi[t][k];
            if (refs[i].c[k] < clo[k]) clo[k] = refs[i].c[k];
            if (refs[i].c[k] > chi[k]) chi[k] = refs[i].c[k];
        }
    }
    bnode *nd = &nv->nodes[node];
    memcpy(nd->lo, lo, sizeof lo);
    memcpy(nd->hi, hi, sizeof hi);
    uint32_t n = end - begin;
    if (n <= 4) { nd->left_or_first = (int32_t)begin; nd->count = (int32_t)n; return; }
    double best_cost = DBL_MAX; int best_axis = -1, best_split = 0;
    for (int k = 0; k < 3; k++) {
        double ext = chi[k] - clo[k];
        if (ext <= 0) continue;
        uint32_t cnt[NBINS] = {0};
        double blo[NBINS][3], bhi[NBINS][3];
        for (int b = 0; b < NBINS; b++) for (int j = 0; j < 3; j++) { blo[b][j] = DBL_MAX; bhi[b][j] = -DBL_MAX; }
        for (uint32_t i = begin; i < end; i++) {
            int b = (int)((refs[i].c[k] - clo[k]) / ext * NBINS);
            if (b >= NBINS) b = NBINS - 1;
            if (b < 0) b = 0;
            cnt[b]++;
            uint32_t t = refs[i].tri;
            for (int j = 0; j < 3; j++) {
                if (tlo[t][j] < blo[b][j]) blo[b][j] = tlo[t][j];
                if (thi[t][j] > bhi[b][j]) bhi[b][j] = thi[t][j];
            }
        }
        double rarea[NBINS]; uint32_t rcnt[NBINS];
        double alo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, ahi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
        uint32_t acc = 0;
        for (int b = NBINS - 1; b > 0; b--) {
            for (int j = 0; j < 3; j++) { if (blo[b][j] < alo[j]) alo[j] = blo[b][j]; if (bhi[b][j] > ahi[j]) ahi[j] = bhi[b][j]; }
            acc += cnt[b];
            rarea[b] = half_area(alo, ahi); rcnt[b] = acc;
        }
        for (int j = 0; j < 3; j++) { alo[j] = DBL_MAX; ahi[j] = -DBL_MAX; }
        acc = 0;
        for (int b = 0; b < NBINS - 1; b++) {
            for (int j = 0; j < 3; j++) { if (blo[b][j] < alo[j]) alo[j] = blo[b][j]; if (bhi[b][j] > ahi[j]) ahi[j] = bhi[b][j]; }
            acc += cnt[b];
            if (acc == 0 || rcnt[b + 1] == 0) continue;
            double cost = half_area(alo, ahi) * acc + rarea[b + 1] * rcnt[b + 1];
            if (cost < best_cost) { best_cost = cost; best_axis = k; best_split = b; }
        }
    }
    uint32_t mid;
    if (best_axis < 0) {
        mid = begin + n / 2; /* all centroids coincide: split by count */
    } else {
        double ext = chi[best_axis] - clo[best_axis];
        uint32_t i = begin, j = end;
        while (i < j) {
            int b = (int)((refs[i].c[best_axis] - clo[best_axis]) / ext * NBINS);
            if (b >= NBINS) b = NBINS - 1;
            if (b <= best_split) i++;
            else { bref tmp = refs[i]; refs[i] = refs[--j]; refs[j] = tmp; }
        }
        mid = i;
        if (mid == begin || mid == end) mid = begin + n / 2;
    }
    uint32_t l = push_node(nv);
    uint32_t r = push_node(nv);
    nv->nodes[node].left_or_first = (int32_t)l;
    nv->nodes[node].count = 0;
    build_rec(s, refs, tlo, thi, begin, mid, nv, l);
    build_rec(s, refs, tlo, thi, mid, end, nv, r);
}

static int build_bvh(orc_scene *s) {
    uint32_t n = s->n_tris;
    bref *refs = (bref *)malloc(sizeof(bref) * (n ? n : 1));
    double (*tlo)[3] = malloc(sizeof(double[3]) * (n ? n : 1));
    double (*thi)[3] = malloc(sizeof(double[3]) * (n ? n : 1));
    if (!refs || !tlo || !thi) return 1;
    for (uint32_t t = 0; t < n; t++) {
        box_tri(s, t, tlo[t], thi[t]);
        for (int k = 0; k < 3; k++) refs[t].c[k] = 0.5 * (tlo[t][k] + thi[t][k]);
        refs[t].tri = t;
    }
    nodevec nv = {0};
    uint32_t root = push_node(&nv);
    if (n == 0) {
        for (int k = 0; k < 3; k++) { nv.nodes[root].lo[k] = 1; nv.nodes[root].hi[k] = -1; }
        nv.nodes[root].left_or_first = 0; nv.nodes[root].count = 0;
        /* count 0 on the root with no children: traversal treats an empty box as a miss */
    } else {
        build_rec(s, refs, tlo, thi, 0, n, &nv, root);
    }
    s->nodes = nv.nodes; s->n_nodes = nv.n;
    s->tri_order = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; i++) s->tri_order[i] = refs[i].tri;
    free(refs); free(tlo); free(thi);
    return 0;
}

/* ------------------------------------------------------------------ ray / triangle (FP64) */
typedef struct { double o[3], d[3], tmin, tmax; int kx, ky, kz; double Sx, Sy, Sz; double inv[3]; } fray;
static void ray_setup(fray *r, const float o[3], const float d[3], float tmin, float tmax) {
    for (int k = 0; k < 3; k++) { r->o[k] = o[k]; r->d[k] = d[k]; }
    r->tmin = tmin; r->tmax = tmax;
    int kz = 0;
    if (fabs(r->d[1]) > fabs(r->d[kz])) kz = 1;
    if (fabs(r->d[2]) > fabs(r->d[kz])) kz = 2;
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (r->d[kz] < 0) { int t = kx; kx = ky; ky = t; }
    r->kx = kx; r->ky = ky; r->kz = kz;
    r->Sx = r->d[kx] / r->d[kz]; r->Sy = r->d[ky] / r->d[kz]; r->Sz = 1.0 / r->d[kz];
    for (int k = 0; k < 3; k++) r->inv[k] = 1.0 / r->d[k];
}
/* Watertight test (Woop, Benthin, Wald, JCGT 2013) in FP64, double-sided.
 * Returns 1 on a hit in (tmin, tmax]; always writes the barycentric margin. */
static int tri_isect(const fray *r, v3 p0, v3 p1, v3 p2, double *t_out, double *margin, double *t_any) {
    double A[3] = {p0.x - r->o[0], p0.y - r->o[1], p0.z - r->o[2]};
    double B[3] = {p1.x - r->o[0], p1.y - r->o[1], p1.z - r->o[2]};
    double C[3] = {p2.x - r->o[0], p2.y - r->o[1], p2.z - r->o[2]};
    int kx = r->kx, ky = r->ky, kz = r->kz;
    double Ax = A[kx] - r->Sx * A[kz], Ay = A[ky] - r->Sy * A[kz];
    double Bx = B[kx] - r->Sx * B[kz], By = B[ky] - r->Sy * B[kz];
    double Cx = C[kx] - r->Sx * C[kz], Cy = C[ky] - r->Sy * C[kz];
    double U = Cx * By - Cy * Bx, Vv = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    double det = U + Vv + W;
    *margin = -1.0; *t_any = INFINITY;
    if (det == 0.0) return 0;
    double Az = r->Sz * A[kz], Bz = r->Sz * B[kz], Cz = r->Sz * C[kz];
    double T = U * Az + Vv * Bz + W * Cz;
    double t = T / det;
    double m = fmin(fmin(U / det, Vv / det), W / det);
    *margin = m;
    *t_any = t;
    if ((U < 0 || Vv < 0 || W < 0) && (U > 0 || Vv > 0 || W > 0)) return 0;
    if (!(t > r->tmin && t <= r->tmax)) return 0;
    *t_out = t;
    return 1;
}
static int box_hit(const fray *r, const bnode *n, double tmax) {
    double t0 = r->tmin, t1 = tmax;
    for (int k = 0; k < 3; k++) {
        if (r->d[k] == 0.0) {
            if (r->o[k] < n->lo[k] || r->o[k] > n->hi[k]) return 0;
            continue;
        }
        double a = (n->lo[k] - r->o[k]) * r->inv[k];
        double b = (n->hi[k] - r->o[k]) * r->inv[k];
        if (a > b) { double t = a; a = b; b = t; }
        a -= 1e-9 * fabs(a) + 1e-300; b += 1e-9 * fabs(b) + 1e-300; /* conservative */
        if (a > t0) t0 = a;
        if (b < t1) t1 = b;
        if (t0 > t1) return 0;
    }
    return 1;
}
static inline v3 tri_v(const orc_scene *s, uint32_t t, int j) { return s->verts[s->tris[3 * t + j]]; }

/* closest hit through the BVH; returns triangle or UINT32_MAX */
static uint32_t bvh_closest(const orc_scene *s, const fray *r, double *t_hit) {
    if (s->n_tris == 0) return UINT32_MAX;
    uint32_t stack[128]; int sp = 0;
    stack[sp++] = 0;
    double best = r->tmax; uint32_t best_tri = UINT32_MAX;
    fray rr = *r;
    while (sp) {
        const bnode *n = &s->nodes[stack[--sp]];
        rr.tmax = best;
        if (!box_hit(&rr, n, best)) continue;
        if (n->count > 0) {
            for (int i = 0; i < n->count; i++) {
                uint32_t t = s->tri_order[n->left_or_first + i];
                double th, m, ta;
                if (tri_isect(&rr, tri_v(s, t, 0), tri_v(s, t, 1), tri_v(s, t, 2), &th, &m, &ta)) {
                    if (th <= best) { best = th; best_tri = t; rr.tmax = best; }
                }
            }
        } else {
            stack[sp++] = (uint32_t)n->left_or_first + 1;
            stack[sp++] = (uint32_t)n->left_or_first;
        }
    }
    *t_hit = best;
    return best_tri;
}
static int bvh_any(const orc_scene *s, const fray *r) {
    if (s->n_tris == 0) return 0;
    uint32_t stack[128]; int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const bnode *n = &s->nodes[stack[--sp]];
        if (!box_hit(r, n, r->tmax)) continue;
        if (n->count > 0) {
            for (int i = 0; i < n->count; i++) {
                uint32_t t = s->tri_order[n->left_or_first + i];
                double th, m, ta;
                if (tri_isect(r, tri_v(s, t, 0), tri_v(s, t, 1), tri_v(s, t, 2), &th, &m, &ta)) return 1;
            }
        } else {
            stack[sp++] = (uint32_t)n->left_or_first + 1;
            stack[sp++] = (uint32_t)n->left_or_first;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ scene create */
int orc_scene_create(const orc_scene_desc *d, orc_scene **out) {
    *out = NULL;
    if (!d) return fail("null desc");
    if (d->camera.width <= 0 || d->camera.height <= 0) return fail("bad resolution");
    orc_scene *s = (orc_scene *)calloc(1, sizeof(orc_scene));
    s->n_verts = d->n_verts; s->n_tris = d->n_tris; s->n_geoms = d->n_geoms;
    s->verts = (v3 *)malloc(sizeof(v3) * (d->n_verts ? d->n_verts : 1));
    for (uint32_t i = 0; i < d->n_verts; i++)
        s->verts[i] = V(d->verts[3 * i], d->verts[3 * i + 1], d->verts[3 * i + 2]);
    s->tris = (uint32_t *)malloc(sizeof(uint32_t) * 3 * (d->n_tris ? d->n_tris : 1));
    memcpy(s->tris, d->tris, sizeof(uint32_t) * 3 * d->n_tris);
    for (uint32_t i = 0; i < 3 * d->n_tris; i++)
        if (s->tris[i] >= d->n_verts) { orc_scene_destroy(s); return fail("triangle index out of range"); }
    s->tri_geom = (uint32_t *)malloc(sizeof(uint32_t) * (d->n_tris ? d->n_tris : 1));
    memcpy(s->tri_geom, d->tri_geom, sizeof(uint32_t) * d->n_tris);
    s->geom_first = (uint32_t *)malloc(sizeof(uint32_t) * (d->n_geoms ? d->n_geoms : 1));
    for (uint32_t g = 0; g < d->n_geoms; g++) s->geom_first[g] = UINT32_MAX;
    for (uint32_t t = 0; t < d->n_tris; t++) {
        uint32_t g = d->tri_geom[t];
        if (g >= d->n_geoms) { orc_scene_destroy(s); return fail("tri_geom out of range"); }
        if (s->geom_first[g] == UINT32_MAX) s->geom_first[g] = t;
    }
    s->geom_material = (int32_t *)malloc(sizeof(int32_t) * (d->n_geoms ? d->n_geoms : 1));
    s->geom_light = (int32_t *)malloc(sizeof(int32_t) * (d->n_geoms ? d->n_geoms : 1));
    memcpy(s->geom_material, d->geom_material, sizeof(int32_t) * d->n_geoms);
    memcpy(s->geom_light, d->geom_light, sizeof(int32_t) * d->n_geoms);
    s->n_mats = d->n_materials;
    s->mats = (bsdf_t *)malloc(sizeof(bsdf_t) * (d->n_materials ? d->n_materials : 1));
    for (uint32_t m = 0; m < d->n_materials; m++) mat_load(&s->mats[m], &d->materials[m]);
    for (uint32_t g = 0; g < d->n_geoms; g++) {
        if (s->geom_light[g] < 0 && (s->geom_material[g] < 0 || (uint32_t)s->geom_material[g] >= s->n_mats)) {
            orc_scene_destroy(s); return fail("geometry without material or light");
        }
    }
    /* lights: DiffuseAreaLight on a TriangleMeshGeometry (diffusearealight.h:21-24,
     * trianglemesh.h:110-129) */
    s->n_lights = d->n_lights;
    s->lights = (light_t *)calloc(d->n_lights, sizeof(light_t));
    double *w = (double *)malloc(sizeof(double) * d->n_lights);
    for (uint32_t l = 0; l < d->n_lights; l++) {
        light_t *L = &s->lights[l];
        L->geom = d->lights[l].geom;
        L->L = RGB(d->lights[l].L[0], d->lights[l].L[1], d->lights[l].L[2]);
        if (L->geom < 0 || (uint32_t)L->geom >= d->n_geoms || s->geom_first[L->geom] == UINT32_MAX) {
            free(w); orc_scene_destroy(s); return fail("light geometry invalid");
        }
        L->first_tri = s->geom_first[L->geom];
        uint32_t c = 0;
        while (L->first_tri + c < d->n_tris && d->tri_geom[L->first_tri + c] == (uint32_t)L->geom) c++;
        L->n_tris = c;
        double *areas = (double *)malloc(sizeof(double) * c);
        double sum = 0;
        for (uint32_t i = 0; i < c; i++) {
            uint32_t t = L->first_tri + i;
            v3 a = tri_v(s, t, 0), b = tri_v(s, t, 1), cc = tri_v(s, t, 2);
            areas[i] = sqrt(vlen2(vcross(vsub(b, a), vsub(cc, a)))) / 2.0;
            sum += areas[i];
        }
        if (!(sum > 0)) { free(areas); free(w); orc_scene_destroy(s); return fail("light with zero area"); }
        L->area = sum; L->inv_area = 1.0 / sum;
        cdf_build(&L->tri_cdf, areas, (int)c);
        free(areas);
        w[l] = sqrt(L->L.r * L->L.r + L->L.g * L->L.g + L->L.b * L->L.b);
    }
    cdf_build(&s->light_cdf, w, (int)d->n_lights);
    s->light_w = w;
    /* camera */
    const orc_camera *c = &d->camera;
    s->W = c->width; s->H = c->height;
    s->cam_origin = V(c->pos[0], c->pos[1], c->pos[2]);
    v3 la = V(c->lookat[0], c->lookat[1], c->lookat[2]), up = V(c->up[0], c->up[1], c->up[2]);
    double fovy = (double)c->fovy_deg * PI / 180.0;
    s->cam_frame.z = vnorm(vsub(la, s->cam_origin));
    s->cam_frame.x = vnorm(vcross(up, s->cam_frame.z));
    s->cam_frame.y = vnorm(vcross(s->cam_frame.z, s->cam_frame.x));
    s->cam_d = (double)s->H * 0.5 / tan(fovy * 0.5);
    s->per_pixel_norm = s->cam_d * s->cam_d / ((double)s->W * (double)s->H);
    s->per_film_norm = s->cam_d * s->cam_d;
    s->aperture = c->aperture_radius;
    s->focal = c->focal_distance;
    s->lens_area = PI * s->aperture * s->aperture;
    if (!(s->focal > 0)) { orc_scene_destroy(s); return fail("focal distance must be > 0"); }
    if (build_bvh(s)) { orc_scene_destroy(s); return fail("bvh build failed"); }
    *out = s;
    return 0;
}
void orc_scene_destroy(orc_scene *s) {
    if (!s) return;
    free(s->verts); free(s->tris); free(s->tri_geom); free(s->geom_first);
    free(s->geom_material); free(s->geom_light); free(s->mats);
    if (s->lights) for (uint32_t l = 0; l < s->n_lights; l++) cdf_free(&s->lights[l].tri_cdf);
    free(s->lights);
    cdf_free(&s->light_cdf);
    free(s->light_w);
    free(s->env_rgb);
    cdf_free(&s->env_cdf);
    free(s->nodes); free(s->tri_order);
    free(s);
}

/* ------------------------------------------------------------------ envmap (light/envmap.h, equirectangular.h) */
#define ENV_CDF_WEIGHT 1.23 /* EnvmapLight::cdfWeight (envmap.h:22-26) */
/* Mapping::PanoramaFromWorld / WorldFromPanorama (mapping.h:91-96,302-323) */
static void panorama_from_world(v3 d, double *u, double *v) {
    *u = atan2(-d.z, -d.x) * INV_PI * 0.5 + 0.5;
    *v = acos(fmax(-1.0, fmin(1.0, d.y))) * INV_PI;
}
static v3 world_from_panorama(double u, double v) {
    double th = u * PI * 2.0, ph = v * PI;
    return V(cos(th) * sin(ph), cos(ph), sin(th) * sin(ph));
}
/* Fimage::evalTexel with S repeat, T clamp (floatimage.h:91-112) */
static rgb env_texel(const orc_scene *s, int x, int y) {
    x %= s->env_w; if (x < 0) x += s->env_w;
    if (y < 0) y = 0;
    if (y > s->env_h - 1) y = s->env_h - 1;
    const double *t = s->env_rgb + 3 * ((size_t)y * s->env_w + x);
    return RGB(t[0], t[1], t[2]);
}
/* EquiRectRep::eval = Fimage::evalBilinear (floatimage.h:120-133) */
static rgb env_eval(const orc_scene *s, v3 d) {
    double u, v;
    panorama_from_world(d, &u, &v);
    double sx = u * s->env_w - 0.5, sy = v * s->env_h - 0.5;
    int x = (int)floor(sx), y = (int)floor(sy);
    double dx1 = sx - x, dy1 = sy - y, dx2 = 1.0 - dx1, dy2 = 1.0 - dy1;
    rgb r = cscale(env_texel(s, x, y), dx2 * dy2);
    r = cadd(r, cscale(env_texel(s, x, y + 1), dx2 * dy1));
    r = cadd(r, cscale(env_texel(s, x + 1, y), dx1 * dy2));
    return cadd(r, cscale(env_texel(s, x + 1, y + 1), dx1 * dy1));
}
/* EquiRectRep::pdfA: solid-angle pdf of sampling direction d (equirectangular.h:64-71) */
static double env_pdf(const orc_scene *s, v3 d) {
    double u, v;
    panorama_from_world(d, &u, &v);
    int x = (int)(u * s->env_w), y = (int)(v * s->env_h);
    if (x > s->env_w - 1) x = s->env_w - 1;
    if (y > s->env_h - 1) y = s->env_h - 1;
    double sine = sqrt(fmax(0.0, 1.0 - d.y * d.y));
    if (sine <= SMALL_VALUE) return 0.0;
    return s->env_cdf.pdf[(size_t)y * s->env_w + x] * s->env_w * s->env_h / sine * INV_PI * INV_PI * 0.5;
}
/* EquiRectRep::sample (equirectangular.h:85-117): texel by the CDF, then
 * (ox, oy) inside it; returns eval / pdf and the pdf */
static rgb env_sample(const orc_scene *s, double u_texel, double ox, double oy, v3 *dir, double *pdf) {
    double pt;
    int idx = cdf_sample(&s->env_cdf, u_texel, &pt, NULL);
    int x = idx % s->env_w, y = idx / s->env_w;
    *dir = world_from_panorama((x + ox) / s->env_w, (y + oy) / s->env_h);
    double sine = sqrt(fmax(0.0, 1.0 - dir->y * dir->y));
    *pdf = sine <= SMALL_VALUE ? 0.0 : pt * s->env_w * s->env_h / sine * INV_PI * INV_PI * 0.5;
    return *pdf > 0.0 ? cscale(env_eval(s, *dir), 1.0 / *pdf) : BLACK;
}
static void rebuild_light_cdf(orc_scene *s) {
    cdf_free(&s->light_cdf);
    int n = (int)s->n_lights + (s->env_on ? 1 : 0);
    double *w = (double *)malloc(sizeof(double) * (n ? n : 1));
    for (uint32_t i = 0; i < s->n_lights; i++) w[i] = s->light_w[i];
    if (s->env_on) w[n - 1] = ENV_CDF_WEIGHT;
    cdf_build(&s->light_cdf, w, n);
    free(w);
}
int orc_scene_set_envmap(orc_scene *s, const float *rgb, int w, int h) {
    free(s->env_rgb); s->env_rgb = NULL;
    cdf_free(&s->env_cdf); memset(&s->env_cdf, 0, sizeof s->env_cdf);
    s->env_on = 0;
    if (rgb && w > 0) {
        if (h <= 0 || w != 2 * h) return fail("envmap has wrong size"); /* equirectangular.h:31 */
        size_t n = (size_t)w * h;
        s->env_rgb = (double *)malloc(sizeof(double) * 3 * n);
        double *wt = (double *)malloc(sizeof(double) * n);
        double sum = 0.0;
        for (size_t i = 0; i < 3 * n; i++) {
            if (!(rgb[i] >= 0.0f) || !isfinite(rgb[i])) { free(wt); return fail("envmap texels must be finite and >= 0"); }
            s->env_rgb[i] = rgb[i];
        }
        for (int y = 0; y < h; y++) {
            double sine = sin((y + 0.5) / h * PI);
            for (int x = 0; x < w; x++) {
                const double *t = s->env_rgb + 3 * ((size_t)y * w + x);
                wt[(size_t)y * w + x] = sine * sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
                sum += wt[(size_t)y * w + x];
            }
        }
        if (!(sum > 0.0)) { free(wt); return fail("envmap is black"); }
        cdf_build(&s->env_cdf, wt, (int)n);
        free(wt);
        s->env_w = w; s->env_h = h; s->env_on = 1;
    }
    rebuild_light_cdf(s);
    return 0;
}
void orc_env_eval(const orc_scene *s, const double d[3], double out[3], double *pdf) {
    v3 v = V(d[0], d[1], d[2]);
    rgb e = env_eval(s, v);
    out[0] = e.r; out[1] = e.g; out[2] = e.b;
    *pdf = env_pdf(s, v);
}
/* EquiRectRep::sample as the reference draws it: the in-texel u is the remapped selection number */
void orc_env_sample_ref(const orc_scene *s, const double u[2], double dir[3], double *pdf, double value[3]) {
    double r;
    cdf_sample(&s->env_cdf, u[0], NULL, &r);
    v3 d;
    rgb v = env_sample(s, u[0], r, u[1], &d, pdf);
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
    value[0] = v.r; value[1] = v.g; value[2] = v.b;
}

/* ------------------------------------------------------------------ light + camera sampling */
/* env != 0: an envmap vertex; pos holds its direction (path.h:126-131) */
typedef struct { v3 pos, n; frame3 f; rgb contrib; double pdfA; int light, env; } lvert;
/* createLightSubpath with numVertices = 1 (generalsubpathsampler.h:185-255),
 * DiffuseAreaLight::sampleLe0ByPdf (diffusearealight.h:53-72),
 * TriangleMeshGeometry::sample (trianglemesh.h:131-159) */
static void sample_light(const orc_scene *s, double u0, double u1, uint32_t key, lvert *lv) {
    double pdfLight, ur;
    int li = cdf_sample(&s->light_cdf, u0, &pdfLight, &ur);
    lv->env = 0;
    if (li == (int)s->n_lights) { /* EnvmapLight::sampleLe0ByPdf (envmap.h:61-77) */
        double pdfW;
        v3 d;
        rgb v = env_sample(s, ur, u01(key, DIM_ENV), u01(key, DIM_ENV + 1), &d, &pdfW);
        lv->pos = d; lv->n = V(0, 0, 0);
        lv->contrib = cscale(v, 1.0 / pdfLight);
        lv->pdfA = pdfLight * pdfW;
        lv->light = li; lv->env = 1;
        return;
    }
    const light_t *L = &s->lights[li];
    double ur2;
    int ti = cdf_sample(&L->tri_cdf, ur, NULL, &ur2);
    double b1, b2;
    barycentric_from_square(ur2, u1, &b1, &b2);
    uint32_t t = L->first_tri + (uint32_t)ti;
    v3 p0 = tri_v(s, t, 0), p1 = tri_v(s, t, 1), p2 = tri_v(s, t, 2);
    lv->pos = vadd(vadd(vmul(p0, 1.0 - b1 - b2), vmul(p1, b1)), vmul(p2, b2));
    lv->n = vnorm(vcross(vsub(p1, p0), vsub(p2, p0)));
    lv->f = make_frame(lv->n);
    double pdfPos = L->inv_area;
    lv->contrib = cscale(cscale(L->L, PI / pdfPos), 1.0 / pdfLight);
    lv->pdfA = pdfLight * pdfPos;
    lv->light = li;
}
int orc_light_sample(const orc_scene *s, const double u[2], double pos[3], double normal[3], double contrib[3], double *pdfA) {
    lvert lv;
    sample_light(s, u[0], u[1], 0u, &lv);
    pos[0] = lv.pos.x; pos[1] = lv.pos.y; pos[2] = lv.pos.z;
    normal[0] = lv.n.x; normal[1] = lv.n.y; normal[2] = lv.n.z;
    contrib[0] = lv.contrib.r; contrib[1] = lv.contrib.g; contrib[2] = lv.contrib.b;
    *pdfA = lv.pdfA;
    return lv.light;
}
/* ThinlensCamera::sampleWe0ByPdf (thinlens.h:121-136) */
static v3 camera_position(const orc_scene *s, double u0, double u1, v3 *lens_local) {
    double dx, dy;
    disk_from_square(u0, u1, &dx, &dy);
    v3 ll = V(dx * s->aperture, dy * s->aperture, 0.0);
    *lens_local = ll;
    return vadd(to_world(&s->cam_frame, ll), s->cam_origin);
}
/* samplePerPixelWe1CosByPdf -> sampleWe1CosByPdf (thinlens.h:138-166, camera.h:19-25) */
static v3 camera_direction(const orc_scene *s, v3 lens_local, int px, int py, double j0, double j1,
                           double raster[2], double *pdfW) {
    double invW = 1.0 / (double)s->W, invH = 1.0 / (double)s->H;
    double ns0 = ((double)px + j0) * invW, ns1 = ((double)py + j1) * invH;
    raster[0] = ns0 * s->W; raster[1] = ns1 * s->H;
    v3 film = V((ns0 - 0.5) * s->W, (ns1 - 0.5) * s->H, -s->cam_d);
    v3 focal = vmul(film, s->focal / film.z);
    v3 dl = vnorm(vsub(focal, lens_local));
    double c = dl.z;
    *pdfW = s->per_film_norm / (c * c * c) * (invW * invH);
    return to_world(&s->cam_frame, dl);
}
void orc_camera_ray(const orc_scene *s, int px, int py, const double ul[2], const double up[2],
                    double org[3], double dir[3], double raster[2], double *pdfW) {
    v3 ll;
    v3 o = camera_position(s, ul[0], ul[1], &ll);
    v3 d = camera_direction(s, ll, px, py, up[0], up[1], raster, pdfW);
    org[0] = o.x; org[1] = o.y; org[2] = o.z;
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
}
/* ThinlensCamera::evalWe1 (thinlens.h:62-80); returns 0 outside the film */
static double camera_we1(const orc_scene *s, v3 cam_pos, v3 w, double raster[2]) {
    v3 pos = to_local(&s->cam_frame, vsub(cam_pos, s->cam_origin));
    v3 dir = to_local(&s->cam_frame, w);
    double c = dir.z;
    if (vdot(w, s->cam_frame.z) < 0.0) return 0.0;
    v3 fp = vadd(vmul(vmul(dir, 1.0 / c), s->focal), pos);
    v3 film = vmul(vmul(fp, -1.0 / s->focal), s->cam_d);
    double rx = (film.x / s->W + 0.5) * s->W;
    double ry = (film.y / s->H + 0.5) * s->H;
    if (!(rx >= 0.0 && rx < s->W && ry >= 0.0 && ry < s->H)) return 0.0; /* bound.h:32-38 */
    raster[0] = rx; raster[1] = ry;
    double c2 = c * c;
    return s->per_film_norm / (c2 * c2);
}
double orc_camera_we1(const orc_scene *s, const double cp[3], const double w[3], double raster[2]) {
    return camera_we1(s, V(cp[0], cp[1], cp[2]), V(w[0], w[1], w[2]), raster);
}

/* ------------------------------------------------------------------ queries in FP32 (embreeaccel.h:336-422) */
typedef struct { uint64_t ext, shd; } counters;
static int occluded(const orc_scene *s, v3 o, v3 d, double tmin, double tmax, counters *cn) {
    float of[3] = {(float)o.x, (float)o.y, (float)o.z}, df[3] = {(float)d.x, (float)d.y, (float)d.z};
    fray r;
    ray_setup(&r, of, df, (float)tmin, (float)tmax);
    cn->shd++;
    return bvh_any(s, &r);
}
/* returns triangle index or UINT32_MAX; writes the FP32-rounded t as the reference does */
static uint32_t closest(const orc_scene *s, v3 o, v3 d, double *t, counters *cn) {
    float of[3] = {(float)o.x, (float)o.y, (float)o.z}, df[3] = {(float)d.x, (float)d.y, (float)d.z};
    fray r;
    ray_setup(&r, of, df, (float)RAY_EPS, INFINITY);
    cn->ext++;
    double th;
    uint32_t tri = bvh_closest(s, &r, &th);
    if (tri != UINT32_MAX) *t = (double)(float)th;
    return tri;
}

/* ------------------------------------------------------------------ medium */
/* HomogeneousMedium (homogeneous.h:11-66): transmittance exp(-sigma_t d) per
 * channel; free-flight distances sampled on one channel chosen uniformly,
 * with the channel-averaged density (homogeneous.h:36-50): sigma_t Tr when the
 * flight ends in the medium, Tr when it passes the segment. */
static inline rgb tr_rgb(const orc_scene *s, double d) {
    return RGB(exp(-s->sig_t[0] * d), exp(-s->sig_t[1] * d), exp(-s->sig_t[2] * d));
}
static inline double pass_pdf(const orc_scene *s, double d) {
    if (!s->has_medium) return 1.0;
    rgb t = tr_rgb(s, d);
    return (t.r + t.g + t.b) / 3.0;
}
static inline double scatter_pdf(const orc_scene *s, double d) {
    rgb t = tr_rgb(s, d);
    return (s->sig_t[0] * t.r + s->sig_t[1] * t.g + s->sig_t[2] * t.b) / 3.0;
}
/* Henyey-Greenstein (1941): p(cos) = (1 - g^2) / (4 pi (1 + g^2 - 2 g cos)^{3/2}), cos
 * between the propagation direction and the scattered direction (g > 0: forward).
 * The reference's HenyeyGreenstein cannot serve (henyengreenstein.h:21-51 throws in
 * evalPdfW and never sets the sampling pdf); this is the textbook function. */
static double hg_pdf(double g, double c) {
    double d = 1.0 + g * g - 2.0 * g * c;
    return (1.0 - g * g) / (4.0 * PI * d * sqrt(d));
}
/* inversion of the HG CDF: cos = (1 + g^2 - s^2) / (2 g), s = (1 - g^2) / (1 - g + 2 g u) */
static v3 hg_sample(double g, v3 fwd, double u0, double u1) {
    double c;
    if (fabs(g) < 1e-3) c = 1.0 - 2.0 * u0;
    else {
        double q = (1.0 - g * g) / (1.0 - g + 2.0 * g * u0);
        c = (1.0 + g * g - q * q) / (2.0 * g);
    }
    if (c > 1.0) c = 1.0;
    if (c < -1.0) c = -1.0;
    double st = sqrt(fmax(0.0, 1.0 - c * c));
    double phi = 2.0 * PI * u1;
    frame3 f = make_frame(fwd);
    return vnorm(to_world(&f, V(st * cos(phi), c, st * sin(phi))));
}
int orc_scene_set_medium(orc_scene *s, const double sigma_a[3], const double sigma_s[3], double g) {
    if (!s) return fail("null scene");
    if (!(g > -1.0 && g < 1.0)) return fail("HG g must be in (-1, 1)");
    int any = 0;
    for (int c = 0; c < 3; c++) {
        if (!(sigma_a[c] >= 0.0) || !(sigma_s[c] >= 0.0)) return fail("negative medium coefficient");
        s->sig_s[c] = sigma_s[c];
        s->sig_t[c] = sigma_a[c] + sigma_s[c];
        any |= s->sig_t[c] > 0.0;
    }
    s->hg_g = g;
    s->has_medium = any;
    return 0;
}
double orc_hg_pdf(double g, double cos_theta) { return hg_pdf(g, cos_theta); }
void orc_hg_sample(double g, const double fwd[3], const double u[2], double out[3]) {
    v3 d = hg_sample(g, V(fwd[0], fwd[1], fwd[2]), u[0], u[1]);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

/* ------------------------------------------------------------------ the estimator */
/* pathnextevent.h:41-150 with the random walk of generalsubpathsampler.h:13-160.
 * base[3] receives splats[0].mContrib; sp_px, sp_py and sp[3] receive splats[1].
 * nee == 0: PrimitivePathRenderer::evalContribution (primitivepath.h:42-58) on
 * the same walk -- no light->camera connection, no NEE, every emitter hit
 * counts with weight 1 (its callback returns mPathContrib * Le). */
static void eval_contribution(const orc_scene *s, int N, int use_mis, int nee, uint32_t key, int px, int py,
                              double base[3], int *sp_ok, int *sp_px, int *sp_py, double sp[3], counters *cn) {
    base[0] = base[1] = base[2] = 0.0;
    *sp_ok = 0;
    /* light subpath: one vertex */
    lvert l;
    memset(&l, 0, sizeof l);
    if (s->light_cdf.n > 0) sample_light(s, u01(key, DIM_LIGHT), u01(key, DIM_LIGHT + 1), key, &l); /* primitive paths never use it */
    double lPdfA = use_mis ? l.pdfA : 1.0;
    /* camera vertex */
    v3 lens_local;
    v3 cpos = camera_position(s, u01(key, DIM_LENS), u01(key, DIM_LENS + 1), &lens_local);
    v3 cn_ = s->cam_frame.z; /* camera geometry normal (thinlens.h:130) */
    /* t = 1: light -> camera connection (pathnextevent.h:75-99) */
    if (nee && l.env) {
        /* envmap vertex: the camera looks along its direction; G = |cos| at the
         * camera only, Le1 = 1, visibility to infinity (path.h:64-76, scene.h:98-101) */
        v3 d = l.pos;
        double G = fabs(vdot(cn_, d));
        double raster[2];
        double we1 = G != 0.0 ? camera_we1(s, cpos, d, raster) : 0.0;
        if (we1 != 0.0 && !occluded(s, cpos, d, RAY_EPS, INFINITY, cn)) {
            double mis = 1.0;
            if (use_mis) {
                v3 dl = to_local(&s->cam_frame, d);
                double camPdfA = s->per_pixel_norm / (dl.z * dl.z * dl.z);
                mis = l.pdfA / (l.pdfA + camPdfA);
            }
            double k = mis * we1 * G;
            sp[0] = k * l.contrib.r; sp[1] = k * l.contrib.g; sp[2] = k * l.contrib.b;
            *sp_px = (int)raster[0]; *sp_py = (int)raster[1];
            *sp_ok = 1;
        }
    } else if (nee) {
        v3 uv = vsub(cpos, l.pos);
        double len2 = vlen2(uv);
        v3 nuv = vmul(uv, 1.0 / sqrt(len2));
        double G = fabs(vdot(l.n, nuv)) * fabs(vdot(cn_, nuv)) / len2;
        if (G != 0.0) {
            double raster[2];
            v3 wcam = vnorm(vsub(l.pos, cpos));
            double we1 = camera_we1(s, cpos, wcam, raster);
            if (we1 != 0.0) {
                v3 wl = to_local(&l.f, vnorm(vsub(cpos, l.pos)));
                double le1 = (wl.y < 0.0) ? 0.0 : INV_PI;
                if (le1 != 0.0) {
                    double len = sqrt(len2);
                    int occ = occluded(s, l.pos, vmul(uv, 1.0 / len), RAY_EPS, len - RAY_EPS, cn);
                    if (!occ) {
                        double mis = 1.0;
                        if (use_mis) {
                            /* c0->pdfA(CameraImaginary, l) = perPixel pdfW * conversion to l */
                            v3 dl = to_local(&s->cam_frame, wcam);
                            double c = dl.z;
                            double pdfW = s->per_pixel_norm / (c * c * c);
                            double camPdfA = pdfW * fabs(vdot(l.n, vmul(nuv, -1.0))) / len2 * pass_pdf(s, len);
                            mis = l.pdfA / (l.pdfA + camPdfA);
                        }
                        double k = mis * we1 * G * le1;
                        /* with a medium the connection is attenuated (the reference omits this: the
                         * camera vertex has no medium, generalsubpathsampler.h:302) */
                        rgb tr = s->has_medium ? tr_rgb(s, len) : RGB(1, 1, 1);
                        sp[0] = k * l.contrib.r * tr.r; sp[1] = k * l.contrib.g * tr.g; sp[2] = k * l.contrib.b * tr.b;
                        *sp_px = (int)raster[0]; *sp_py = (int)raster[1];
                        *sp_ok = 1;
                    }
                }
            }
        }
    }
    /* camera direction */
    double raster0[2], pdfW;
    v3 dir = camera_direction(s, lens_local, px, py, u01(key, DIM_PIXEL), u01(key, DIM_PIXEL + 1), raster0, &pdfW);
    /* random walk */
    v3 ro = cpos, rd = dir;
    v3 prev_pos = cpos;
    rgb prev_contrib = RGB(1, 1, 1);  /* contrib1 of samplePerPixelWe1CosByPdf */
    rgb prev_path = RGB(1, 1, 1);     /* camera vertex mPathContrib = We0/pdf = 1 */
    double prevPdfW = use_mis ? pdfW : 0.0;
    int cNum = 1;
    int count = 0;
    for (int bounce = 0;; bounce++) {
        double t = INFINITY;
        uint32_t tri = closest(s, ro, rd, &t, cn);
        /* free flight through the medium (generalsubpathsampler.h:40-53, homogeneous.h:36-50):
         * the 2 medium draws of the bounce pick a channel and a distance */
        int in_medium = 0;
        rgb seg = RGB(1, 1, 1);
        double seg_pdf = 1.0;
        if (s->has_medium) {
            uint32_t dm = DIM_BOUNCE0 + DIMS_PER_BOUNCE * (uint32_t)bounce + BD_MEDIUM;
            double um = u01(key, dm), uc = u01(key, dm + 1);
            int ch = (int)(uc * 3.0);
            if (ch > 2) ch = 2;
            double tf = s->sig_t[ch] > 0.0 ? -log(1.0 - um) / s->sig_t[ch] : INFINITY;
            if (tf < (tri == UINT32_MAX ? INFINITY : t)) {
                in_medium = 1;
                t = tf;
                seg_pdf = scatter_pdf(s, tf);
                seg = cscale(cmul(rgb3(s->sig_s), tr_rgb(s, tf)), 1.0 / seg_pdf);
            } else if (tri != UINT32_MAX) {
                seg_pdf = pass_pdf(s, t);
                seg = cscale(tr_rgb(s, t), 1.0 / seg_pdf);
            }
        }
        if (!in_medium && tri == UINT32_MAX) {
            /* the envmap vertex Scene::intersect returns on a miss (scene.h:82-88): an
             * emitter hit (pathnextevent.h:128-146) with Le1 = 1, Le0 = the lookup, and
             * area pdfs equal to the solid-angle ones (path.h:47-60) */
            if (s->env_on) {
                rgb path = cmul(prev_contrib, prev_path);
                double mis = nee ? 0.0 : 1.0;
                if (use_mis) {
                    double lightPdfA = s->light_cdf.pdf[s->n_lights] * env_pdf(s, rd);
                    mis = prevPdfW / (lightPdfA + prevPdfW);
                }
                rgb Le = env_eval(s, rd);
                base[0] += mis * path.r * Le.r; base[1] += mis * path.g * Le.g; base[2] += mis * path.b * Le.b;
            }
            break; /* without an envmap: callback(nullptr) -> stop */
        }
        v3 p = vadd(ro, vmul(rd, t));
        rgb path = cmul(cmul(prev_contrib, prev_path), seg);
        cNum++;
        if (in_medium) {
            /* medium vertex: NEE through the phase function (no cosine at this vertex) */
            if (cNum == N) break;
            v3 uv = vsub(l.pos, p);
            double len2 = vlen2(uv);
            double len = sqrt(len2);
            v3 nuv = vmul(uv, 1.0 / len);
            double G = fabs(vdot(l.n, nuv)) / len2;
            if (nee && G != 0.0) {
                int occ = occluded(s, p, nuv, RAY_EPS, len - RAY_EPS, cn);
                if (!occ) {
                    double ph = hg_pdf(s->hg_g, vdot(rd, nuv));
                    v3 wl = to_local(&l.f, vnorm(vsub(p, l.pos)));
                    double le1 = (wl.y < 0.0) ? 0.0 : INV_PI;
                    if (le1 != 0.0) {
                        double mis = 1.0;
                        if (use_mis) {
                            double phPdfA = ph * fabs(vdot(l.n, nuv)) / len2 * pass_pdf(s, len);
                            mis = lPdfA / (lPdfA + phPdfA);
                        }
                        rgb c = cmul(cmul(path, l.contrib), tr_rgb(s, len));
                        double k = mis * le1 * G * ph;
                        base[0] += k * c.r; base[1] += k * c.g; base[2] += k * c.b;
                    }
                }
            }
            if (++count == N - 1) break;
            /* phase sampling: HG is sampled exactly, weight 1 */
            uint32_t dim = DIM_BOUNCE0 + DIMS_PER_BOUNCE * (uint32_t)bounce + BD_BSDF;
            v3 nd = hg_sample(s->hg_g, rd, u01(key, dim), u01(key, dim + 1));
            prevPdfW = use_mis ? hg_pdf(s->hg_g, vdot(rd, nd)) : 0.0;
            prev_contrib = RGB(1, 1, 1);
            prev_path = path;
            prev_pos = p;
            ro = p;
            rd = nd;
            continue;
        }
        /* hit vertex (embreeaccel.h:398-420) */
        v3 p0 = tri_v(s, tri, 0), p1 = tri_v(s, tri, 1), p2 = tri_v(s, tri, 2);
        v3 ng = vnorm(vcross(vsub(p1, p0), vsub(p2, p0)));
        frame3 f = make_frame(ng);
        uint32_t g = s->tri_geom[tri];
        /* area pdf of the sampling that produced this vertex (path.h:47-60) */
        double cPdfA = prevPdfW;
        if (use_mis) {
            v3 uv = vsub(p, prev_pos);
            double len2 = vlen2(uv);
            v3 nuv = vmul(uv, 1.0 / sqrt(len2));
            cPdfA *= fabs(vdot(ng, nuv)) / len2 * seg_pdf;
        }
        int is_light = s->geom_light[g] >= 0;
        if (!is_light && cNum != N && l.env) {
            /* NEE toward an envmap vertex: G = |cos| at p, visibility to infinity */
            v3 nuv = l.pos;
            double G = fabs(vdot(ng, nuv));
            if (nee && G != 0.0 && !occluded(s, p, nuv, RAY_EPS, INFINITY, cn)) {
                const bsdf_t *b = &s->mats[s->geom_material[g]];
                v3 wi = to_local(&f, vnorm(vsub(prev_pos, p)));
                v3 wo = to_local(&f, nuv);
                rgb fs = bsdf_eval(b, wi, wo);
                if (!czero(fs)) {
                    double mis = 1.0;
                    if (use_mis) mis = lPdfA / (lPdfA + bsdf_pdf(b, wi, wo));
                    rgb c = cmul(cmul(path, fs), l.contrib);
                    base[0] += mis * G * c.r; base[1] += mis * G * c.g; base[2] += mis * G * c.b;
                }
            }
        } else if (!is_light && cNum != N) {
            /* NEE (pathnextevent.h:106-127) */
            v3 uv = vsub(l.pos, p);
            double len2 = vlen2(uv);
            double len = sqrt(len2);
            v3 nuv = vmul(uv, 1.0 / len);
            double G = fabs(vdot(ng, nuv)) * fabs(vdot(l.n, nuv)) / len2;
            if (nee && G != 0.0) {
                int occ = occluded(s, p, nuv, RAY_EPS, len - RAY_EPS, cn);
                if (!occ) {
                    const bsdf_t *b = &s->mats[s->geom_material[g]];
                    v3 wi = to_local(&f, vnorm(vsub(prev_pos, p)));
                    v3 wo = to_local(&f, vnorm(vsub(l.pos, p)));
                    rgb fs = bsdf_eval(b, wi, wo);
                    if (!czero(fs)) {
                        v3 wl = to_local(&l.f, vnorm(vsub(p, l.pos)));
                        double le1 = (wl.y < 0.0) ? 0.0 : INV_PI;
                        if (le1 != 0.0) {
                            double mis = 1.0;
                            if (use_mis) {
                                double bsdfPdfA = bsdf_pdf(b, wi, wo) * fabs(vdot(l.n, nuv)) / len2 * pass_pdf(s, len);
                                mis = lPdfA / (lPdfA + bsdfPdfA);
                            }
                            rgb c = cmul(cmul(path, fs), l.contrib);
                            if (s->has_medium) c = cmul(c, tr_rgb(s, len));
                            double k = mis * le1 * G;
                            base[0] += k * c.r; base[1] += k * c.g; base[2] += k * c.b;
                        }
                    }
                }
            }
        } else if (is_light) {
            /* emitter hit (pathnextevent.h:128-146) */
            const light_t *L = &s->lights[s->geom_light[g]];
            double mis = nee ? 0.0 : 1.0;
            if (use_mis) {
                double lightPdfA = s->light_cdf.pdf[s->geom_light[g]] * L->inv_area;
                mis = cPdfA / (lightPdfA + cPdfA);
            }
            v3 wl = to_local(&f, vnorm(vsub(prev_pos, p)));
            double le1 = (wl.y < 0.0) ? 0.0 : INV_PI;
            double k = mis * le1 * PI;
            base[0] += k * path.r * L->L.r; base[1] += k * path.g * L->L.g; base[2] += k * path.b * L->L.b;
        } else {
            break; /* surface at the last vertex: callback returns false */
        }
        if (++count == N - 1) break;
        if (is_light) break;
        /* BSDF sampling (generalsubpathsampler.h:81-130) */
        const bsdf_t *b = &s->mats[s->geom_material[g]];
        uint32_t dim = DIM_BOUNCE0 + DIMS_PER_BOUNCE * (uint32_t)bounce + BD_BSDF;
        v3 wi = to_local(&f, vneg(rd));
        v3 wo;
        double pdf = 0.0;
        rgb w = bsdf_sample(b, wi, u01(key, dim), u01(key, dim + 1), &wo, &pdf);
        if (czero(w)) break;
        prev_contrib = w;
        prevPdfW = use_mis ? pdf : 0.0;
        prev_path = path;
        prev_pos = p;
        ro = p;
        rd = to_world(&f, wo);
    }
}

/* VisualizeGeometryNormal (visualizegeomnormal.h:45-66): the camera sample's
 * first hit gives (1, 1, 1) on an emitter, else the unit geometric normal
 * normalize((p1 - p0) x (p2 - p0)) (embreeaccel.h:418), unflipped; a miss 0. */
static void visualize_sample(const orc_scene *s, uint32_t key, int px, int py, double out[3], counters *cn) {
    out[0] = out[1] = out[2] = 0.0;
    v3 lens_local;
    v3 cpos = camera_position(s, u01(key, DIM_LENS), u01(key, DIM_LENS + 1), &lens_local);
    double raster[2], pdfW;
    v3 dir = camera_direction(s, lens_local, px, py, u01(key, DIM_PIXEL), u01(key, DIM_PIXEL + 1), raster, &pdfW);
    double t = INFINITY;
    uint32_t tri = closest(s, cpos, dir, &t, cn);
    if (tri == UINT32_MAX) {
        if (s->env_on) out[0] = out[1] = out[2] = 1.0; /* the envmap vertex is a light vertex */
        return;
    }
    if (s->geom_light[s->tri_geom[tri]] >= 0) {
        out[0] = out[1] = out[2] = 1.0;
        return;
    }
    v3 p0 = tri_v(s, tri, 0), p1 = tri_v(s, tri, 1), p2 = tri_v(s, tri, 2);
    v3 ng = vnorm(vcross(vsub(p1, p0), vsub(p2, p0)));
    out[0] = ng.x; out[1] = ng.y; out[2] = ng.z;
}

static void atomic_add_double(double *addr, double v) {
    uint64_t *a = (uint64_t *)addr;
    uint64_t oldb = __atomic_load_n(a, __ATOMIC_RELAXED), newb;
    do {
        double o; memcpy(&o, &oldb, 8);
        double n = o + v;
        memcpy(&newb, &n, 8);
    } while (!__atomic_compare_exchange_n(a, &oldb, newb, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED));
}
/* LightSplattingRenderer (lightsplatting.h:58-108) with createLightSubpath
 * (generalsubpathsampler.h:185-268) and its non-volumetric randomWalk: a light
 * subpath of up to N - 1 vertices; its light vertex and every surface vertex
 * connect to the camera vertex with weight 1, adding
 * c0.pathContrib (1) * We1 * G * visibility * scatter * pathContrib to the
 * splat film at the camera raster.  An emitter hit ends the walk. */
static void splat_add(const orc_scene *s, double *splat, const double raster[2], rgb c) {
    double *dst = splat + 3 * ((size_t)(int)raster[1] * s->W + (size_t)(int)raster[0]);
    atomic_add_double(dst + 0, c.r);
    atomic_add_double(dst + 1, c.g);
    atomic_add_double(dst + 2, c.b);
}
static void light_subpath(const orc_scene *s, int N, uint32_t key, double *splat, counters *cn) {
    if (N < 2) return;
    lvert l;
    sample_light(s, u01(key, DIM_LIGHT), u01(key, DIM_LIGHT + 1), key, &l);
    v3 lens_local;
    v3 cpos = camera_position(s, u01(key, DIM_LENS), u01(key, DIM_LENS + 1), &lens_local);
    v3 cz = s->cam_frame.z;
    /* the light vertex (epCallbackFunc, lightsplatting.h:77-89) */
    {
        v3 uv = vsub(cpos, l.pos);
        double len2 = vlen2(uv), len = sqrt(len2);
        v3 nuv = vmul(uv, 1.0 / len);
        double G = fabs(vdot(l.n, nuv)) * fabs(vdot(cz, nuv)) / len2;
        double raster[2];
        double we1 = G != 0.0 ? camera_we1(s, cpos, vnorm(vsub(l.pos, cpos)), raster) : 0.0;
        v3 wl = to_local(&l.f, nuv);
        double le1 = wl.y < 0.0 ? 0.0 : INV_PI;
        if (we1 != 0.0 && le1 != 0.0 && !occluded(s, l.pos, nuv, RAY_EPS, len - RAY_EPS, cn))
            splat_add(s, splat, raster, cscale(l.contrib, we1 * G * le1));
    }
    if (N - 1 == 1) return;
    /* Le1: cosine-distributed direction in the light's frame, contribution 1 */
    v3 ro = l.pos, rd = to_world(&l.f, cosine_hemisphere(u01(key, DIM_PIXEL), u01(key, DIM_PIXEL + 1)));
    rgb T = l.contrib;
    for (int bounce = 0;; bounce++) {
        const int k = bounce + 2; /* vertex of the light subpath */
        double t = INFINITY;
        uint32_t tri = closest(s, ro, rd, &t, cn);
        if (tri == UINT32_MAX) break;
        uint32_t g = s->tri_geom[tri];
        if (s->geom_light[g] >= 0) break; /* callback adds nothing; the walk ends at a light */
        v3 p = vadd(ro, vmul(rd, t));
        v3 p0 = tri_v(s, tri, 0), p1 = tri_v(s, tri, 1), p2 = tri_v(s, tri, 2);
        v3 ng = vnorm(vcross(vsub(p1, p0), vsub(p2, p0)));
        frame3 f = make_frame(ng);
        const bsdf_t *b = &s->mats[s->geom_material[g]];
        v3 wi = to_local(&f, vneg(rd));
        /* connection to the camera (lightsplatting.h:91-105) */
        v3 uv = vsub(cpos, p);
        double len2 = vlen2(uv), len = sqrt(len2);
        v3 nuv = vmul(uv, 1.0 / len);
        double G = fabs(vdot(ng, nuv)) * fabs(vdot(cz, nuv)) / len2;
        double raster[2];
        double we1 = G != 0.0 ? camera_we1(s, cpos, vnorm(vsub(p, cpos)), raster) : 0.0;
        if (we1 != 0.0) {
            rgb fs = bsdf_eval(b, wi, to_local(&f, nuv));
            if (!czero(fs) && !occluded(s, p, nuv, RAY_EPS, len - RAY_EPS, cn))
                splat_add(s, splat, raster, cscale(cmul(T, fs), we1 * G));
        }
        if (k == N - 1) break;
        uint32_t dim = DIM_BOUNCE0 + DIMS_PER_BOUNCE * (uint32_t)bounce + BD_BSDF;
        v3 wo;
        double pdf = 0.0;
        rgb w = bsdf_sample(b, wi, u01(key, dim), u01(key, dim + 1), &wo, &pdf);
        if (czero(w)) break;
        T = cmul(T, w);
        ro = p;
        rd = to_world(&f, wo);
    }
}

/* ------------------------------------------------------------------ bidirectional path tracing
 * BidirPathRenderer (bidirpath.h:44-421) over GeneralSubpathSampler with
 * PdfA requests and no volumetric sampling (generalsubpathsampler.h:13-160,
 * 185-322; the renderer's sampler is built non-volumetric, so a medium
 * changes nothing).  Vertices are 1-based as in the reference: L[1] is the
 * light point, C[1] the camera vertex.  pf is the area pdf with which a
 * vertex's own subpath made it (the reference's mPdfFromLight on the light
 * side, mPdfFromCamera on the camera side), pr the area pdf with which the
 * other subpath would have made it. */
#define BDPT_MAXV 64
enum { BV_CAMERA = 0, BV_LIGHT = 1, BV_SURFACE = 2 };
typedef struct {
    v3 p, n;     /* position, geometric normal (the camera's: its viewing axis) */
    frame3 f;
    v3 wi;       /* surfaces: unit direction to the previous vertex of its own subpath */
    rgb T;       /* mPathContrib */
    double pf, pr;
    int kind, id; /* id: material (surface) or light index (light) */
} bvert;
/* the light subpath's walk draws from its own stream */
static inline uint32_t bdpt_light_key(uint32_t key) { return mix32(key ^ 0x3c6ef372u); }
/* Vertex::ConvertToAreaFromSolidAngle (path.h:47-60) for surface, light and camera vertices */
static double to_area(v3 from, v3 to, v3 to_n) {
    v3 uv = vsub(to, from);
    double l2 = vlen2(uv);
    return fabs(vdot(to_n, vmul(uv, 1.0 / sqrt(l2)))) / l2;
}
/* Vertex::pdfA(from, to) of vertex x (path.h:101, 197-205, 307-315, 381-390):
 * BSDF pdf, DiffuseAreaLight::evalPdfLe1 (diffusearealight.h:34-38) or the
 * camera's per-pixel pdf (generalsubpathsampler.h:324-327); `from` is unused
 * (the imaginary vertex) for lights and the camera */
static double bv_pdfA(const orc_scene *s, const bvert *x, const bvert *from, const bvert *to) {
    v3 w = vnorm(vsub(to->p, x->p));
    double pw;
    if (x->kind == BV_SURFACE) {
        pw = bsdf_pdf(&s->mats[x->id], to_local(&x->f, vnorm(vsub(from->p, x->p))), to_local(&x->f, w));
    } else if (x->kind == BV_LIGHT) {
        double c = to_local(&x->f, w).y;
        pw = c > 0.0 ? c * INV_PI : 0.0;
    } else {
        double c = to_local(&s->cam_frame, w).z;
        pw = s->per_pixel_norm / (c * c * c);
    }
    return pw * to_area(x->p, to->p, to->n);
}
/* computeMisWeight with power 1 and the ratio form (bidirpath.h:125-194):
 * connected vertex i <= s is L[i], i > s is C[s + t - i + 1]; lp[i] / cp[i]
 * are the area pdfs with which the light / camera side makes vertex i, those
 * at the two vertices each side of the connection re-evaluated for it.
 * Returns 0 (weight 0) unless vertex 1 is on a light. */
static int bdpt_pdf_arrays(const orc_scene *s, const bvert *L, const bvert *C, int sN, int tN, double *lp,
                           double *cp) {
    const int n = sN + tN;
    for (int i = 1; i <= n; i++) {
        const bvert *x = i <= sN ? &L[i] : &C[n - i + 1];
        lp[i] = i <= sN ? x->pf : x->pr;
        cp[i] = i <= sN ? x->pr : x->pf;
    }
    if ((sN >= 1 ? L[1].kind : C[n].kind) != BV_LIGHT) return 0;
    const bvert *c = &C[tN], *cq = tN >= 2 ? &C[tN - 1] : NULL;
    const bvert *l = sN >= 1 ? &L[sN] : NULL, *lq = sN >= 2 ? &L[sN - 1] : NULL;
    if (sN >= 1) cp[sN] = bv_pdfA(s, c, cq, l);
    if (sN >= 2) cp[sN - 1] = bv_pdfA(s, l, c, lq);
    if (sN >= 1 && tN >= 2) lp[sN + 1] = bv_pdfA(s, l, lq, c);
    if (sN >= 1 && tN >= 2) lp[sN + 2] = bv_pdfA(s, c, l, cq);
    return 1;
}
static double bdpt_mis(const orc_scene *s, const bvert *L, const bvert *C, int sN, int tN) {
    const int n = sN + tN;
    double lp[2 * BDPT_MAXV + 3], cp[2 * BDPT_MAXV + 3];
    if (!bdpt_pdf_arrays(s, L, C, sN, tN, lp, cp)) return 0.0;
    double inv = 1.0, r = 1.0;
    for (int i = sN; i > 0; i--) {
        if (lp[i] == 0.0) return 0.0;
        r *= cp[i] / lp[i];
        inv += r;
    }
    r = 1.0;
    for (int i = sN + 1; i < n; i++) {
        if (cp[i] == 0.0) return 0.0;
        r *= lp[i] / cp[i];
        inv += r;
    }
    return 1.0 / inv;
}
/* bdptConnect (bidirpath.h:79-123): s = 0 is the camera subpath ending on an
 * emitter, t = 1 a connection to the camera vertex (raster from We1), else a
 * connection of two surface vertices; visibility from the light side */
static rgb bdpt_connect(const orc_scene *s, const bvert *L, const bvert *C, int sN, int tN, double raster[2],
                        counters *cn) {
    if (sN == 0) {
        const bvert *c = &C[tN], *cq = &C[tN - 1];
        if (c->kind != BV_LIGHT) return BLACK;
        if (to_local(&c->f, vnorm(vsub(cq->p, c->p))).y < 0.0) return BLACK;
        return cscale(cmul(c->T, s->lights[c->id].L), INV_PI * PI); /* Le1 * Le0 */
    }
    const bvert *l = &L[sN], *c = &C[tN];
    if (tN == 1 ? c->kind != BV_CAMERA : c->kind != BV_SURFACE) return BLACK;
    v3 uv = vsub(c->p, l->p);
    double len2 = vlen2(uv), len = sqrt(len2);
    v3 nuv = vmul(uv, 1.0 / len);
    double G = fabs(vdot(l->n, nuv)) * fabs(vdot(c->n, nuv)) / len2;
    if (G == 0.0) return BLACK;
    rgb cs;
    if (tN == 1) {
        double we1 = camera_we1(s, c->p, vneg(nuv), raster);
        if (we1 == 0.0) return BLACK;
        cs = RGB(we1, we1, we1);
    } else {
        cs = bsdf_eval(&s->mats[c->id], to_local(&c->f, c->wi), to_local(&c->f, vneg(nuv)));
        if (czero(cs)) return BLACK;
    }
    rgb ls;
    if (sN == 1) {
        if (to_local(&l->f, nuv).y < 0.0) return BLACK;
        ls = RGB(INV_PI, INV_PI, INV_PI);
    } else {
        ls = bsdf_eval(&s->mats[l->id], to_local(&l->f, l->wi), to_local(&l->f, nuv));
        if (czero(ls)) return BLACK;
    }
    if (occluded(s, l->p, nuv, RAY_EPS, len - RAY_EPS, cn)) return BLACK;
    return cmul(cscale(cmul(l->T, ls), G), cmul(cs, c->T));
}
/* one camera sample (bidirpath.h:364-395): light subpath, camera subpath, all
 * strategies with s + t <= N; t = 1 goes to the splat film, the rest to base.
 * The light subpath stops at N - 1 vertices: the reference traces an N-th one
 * that no strategy uses (s <= N - 1). */
static void bdpt_sample(const orc_scene *s, int N, uint32_t key, int px, int py, double base[3], double *splat,
                        counters *cn, orc_bdpt_record *rec, int max_rec, int *n_rec) {
    bvert L[BDPT_MAXV + 1], C[BDPT_MAXV + 1];
    base[0] = base[1] = base[2] = 0.0;
    /* light subpath (traceLightSubpath, bidirpath.h:196-219) */
    lvert lv;
    sample_light(s, u01(key, DIM_LIGHT), u01(key, DIM_LIGHT + 1), key, &lv);
    memset(&L[1], 0, sizeof L[1]);
    L[1].p = lv.pos; L[1].n = lv.n; L[1].f = lv.f; L[1].T = lv.contrib; L[1].pf = lv.pdfA;
    L[1].kind = BV_LIGHT; L[1].id = lv.light;
    int nL = 1;
    const uint32_t kl = bdpt_light_key(key);
    if (N - 1 >= 2) {
        v3 ro = lv.pos, rd = to_world(&lv.f, cosine_hemisphere(u01(kl, DIM_PIXEL), u01(kl, DIM_PIXEL + 1)));
        double pdfW = fmax(to_local(&lv.f, rd).y, 0.0) * INV_PI; /* sampleLe1CosByPdf, contribution 1 */
        rgb w = RGB(1, 1, 1);
        for (int bounce = 0;; bounce++) {
            double t = INFINITY;
            uint32_t tri = closest(s, ro, rd, &t, cn);
            if (tri == UINT32_MAX) break;
            uint32_t g = s->tri_geom[tri];
            if (s->geom_light[g] >= 0) break; /* the callback keeps surfaces only; a light ends the walk */
            bvert *x = &L[nL + 1], *prev = &L[nL];
            v3 p0 = tri_v(s, tri, 0), p1 = tri_v(s, tri, 1), p2 = tri_v(s, tri, 2);
            x->p = vadd(ro, vmul(rd, t));
            x->n = vnorm(vcross(vsub(p1, p0), vsub(p2, p0)));
            x->f = make_frame(x->n);
            x->wi = vneg(rd);
            x->T = cmul(prev->T, w);
            x->pf = pdfW * to_area(prev->p, x->p, x->n);
            x->pr = 0.0;
            x->kind = BV_SURFACE; x->id = (int)s->geom_material[g];
            nL++;
            if (nL == N - 1) break;
            const bsdf_t *b = &s->mats[x->id];
            uint32_t dim = DIM_BOUNCE0 + DIMS_PER_BOUNCE * (uint32_t)bounce + BD_BSDF;
            v3 wi = to_local(&x->f, x->wi), wo;
            double pdf = 0.0;
            w = bsdf_sample(b, wi, u01(kl, dim), u01(kl, dim + 1), &wo, &pdf);
            if (czero(w)) break;
            /* the previous vertex's pdf from the camera side, given this direction (bidirpath.h:207-212) */
            prev->pr = bsdf_pdf(b, wo, wi) * to_area(x->p, prev->p, prev->n);
            pdfW = pdf;
            ro = x->p;
            rd = to_world(&x->f, wo);
        }
    }
    /* camera subpath (traceCameraSubpath, bidirpath.h:221-257) */
    v3 lens_local;
    memset(&C[1], 0, sizeof C[1]);
    C[1].p = camera_position(s, u01(key, DIM_LENS), u01(key, DIM_LENS + 1), &lens_local);
    C[1].n = s->cam_frame.z;
    C[1].T = RGB(1, 1, 1);
    C[1].kind = BV_CAMERA;
    int nC = 1;
    {
        double raster0[2], pdfW;
        v3 ro = C[1].p;
        v3 rd = camera_direction(s, lens_local, px, py, u01(key, DIM_PIXEL), u01(key, DIM_PIXEL + 1), raster0, &pdfW);
        rgb w = RGB(1, 1, 1);
        for (int bounce = 0; nC < N; bounce++) {
            double t = INFINITY;
            uint32_t tri = closest(s, ro, rd, &t, cn);
            if (tri == UINT32_MAX) break;
            uint32_t g = s->tri_geom[tri];
            bvert *x = &C[nC + 1], *prev = &C[nC];
            v3 p0 = tri_v(s, tri, 0), p1 = tri_v(s, tri, 1), p2 = tri_v(s, tri, 2);
            x->p = vadd(ro, vmul(rd, t));
            x->n = vnorm(vcross(vsub(p1, p0), vsub(p2, p0)));
            x->f = make_frame(x->n);
            x->wi = vneg(rd);
            x->T = cmul(prev->T, w);
            x->pf = pdfW * to_area(prev->p, x->p, x->n);
            nC++;
            if (s->geom_light[g] >= 0) {
                /* emitter: its light-side pdfs (bidirpath.h:240-249) end the walk */
                const int li = s->geom_light[g];
                x->kind = BV_LIGHT; x->id = li;
                x->pr = s->light_cdf.pdf[li] * s->lights[li].inv_area;
                prev->pr = bv_pdfA(s, x, NULL, prev);
                break;
            }
            x->kind = BV_SURFACE; x->id = (int)s->geom_material[g];
            x->pr = 0.0;
            if (nC == N) break;
            const bsdf_t *b = &s->mats[x->id];
            uint32_t dim = DIM_BOUNCE0 + DIMS_PER_BOUNCE * (uint32_t)bounce + BD_BSDF;
            v3 wi = to_local(&x->f, x->wi), wo;
            double pdf = 0.0;
            w = bsdf_sample(b, wi, u01(key, dim), u01(key, dim + 1), &wo, &pdf);
            if (czero(w)) break;
            prev->pr = bsdf_pdf(b, wo, wi) * to_area(x->p, prev->p, prev->n); /* bidirpath.h:234-239 */
            pdfW = pdf;
            ro = x->p;
            rd = to_world(&x->f, wo);
        }
    }
    /* every strategy (evalContribution, bidirpath.h:259-301) */
    for (int sN = 0; sN <= nL; sN++)
        for (int tN = 1; tN <= nC; tN++) {
            if (sN + tN <= 1 || sN + tN > N) continue;
            double raster[2];
            rgb c = bdpt_connect(s, L, C, sN, tN, raster, cn);
            if (czero(c)) continue;
            const double w = bdpt_mis(s, L, C, sN, tN);
            if (rec && *n_rec < max_rec) {
                orc_bdpt_record *r = &rec[(*n_rec)++];
                memset(r, 0, sizeof *r);
                r->s = sN; r->t = tN; r->n = sN + tN; r->weight = w;
                r->contrib[0] = c.r; r->contrib[1] = c.g; r->contrib[2] = c.b;
                double lp[2 * BDPT_MAXV + 3], cp[2 * BDPT_MAXV + 3];
                bdpt_pdf_arrays(s, L, C, sN, tN, lp, cp);
                for (int i = 1; i <= r->n; i++) {
                    const bvert *x = i <= sN ? &L[i] : &C[r->n - i + 1];
                    r->pos[i - 1][0] = x->p.x; r->pos[i - 1][1] = x->p.y; r->pos[i - 1][2] = x->p.z;
                    r->nrm[i - 1][0] = x->n.x; r->nrm[i - 1][1] = x->n.y; r->nrm[i - 1][2] = x->n.z;
                    r->kind[i - 1] = x->kind; r->id[i - 1] = x->id;
                    r->lp[i - 1] = lp[i]; r->cp[i - 1] = cp[i];
                }
            }
            c = cscale(c, w);
            if (tN == 1) {
                if (splat) splat_add(s, splat, raster, c);
            } else {
                base[0] += c.r; base[1] += c.g; base[2] += c.b;
            }
        }
}

/* ------------------------------------------------------------------ render driver (pathnextevent.h:152-191) */
typedef struct {
    const orc_scene *s;
    const orc_render_params *p;
    int mode;  /* 0: PathNeeRenderer, 1: PrimitivePathRenderer, 2: VisualizeGeometryNormal, 3: LightSplattingRenderer,
                  4: BidirPathRenderer */
    double *base, *splat;
    int tiles_x, tiles_y, x0, y0, x1, y1;
    volatile int next_tile;
    counters total;
    pthread_mutex_t mu;
} job_t;
#define TILE 16
static void *worker(void *arg) {
    job_t *j = (job_t *)arg;
    const orc_scene *s = j->s;
    const orc_render_params *p = j->p;
    counters cn = {0, 0};
    int ntiles = j->tiles_x * j->tiles_y;
    for (;;) {
        int tile = __atomic_fetch_add(&j->next_tile, 1, __ATOMIC_RELAXED);
        if (tile >= ntiles) break;
        int tx = tile % j->tiles_x, ty = tile / j->tiles_x;
        int xa = j->x0 + tx * TILE, ya = j->y0 + ty * TILE;
        int xb = xa + TILE < j->x1 ? xa + TILE : j->x1;
        int yb = ya + TILE < j->y1 ? ya + TILE : j->y1;
        for (int py = ya; py < yb; py++)
            for (int px = xa; px < xb; px++) {
                uint32_t pixel = (uint32_t)(py * s->W + px);
                double acc[3] = {0, 0, 0};
                for (int k = 0; k < p->spp; k++) {
                    uint32_t sample = (uint32_t)(p->first_sample + k);
                    uint32_t key = path_key(p->seed, pixel, sample);
                    double b[3], sp[3];
                    int ok, sx, sy;
                    if (j->mode == 2) {
                        visualize_sample(s, key, px, py, b, &cn);
                        acc[0] += b[0]; acc[1] += b[1]; acc[2] += b[2];
                        continue;
                    }
                    if (j->mode == 3) {
                        light_subpath(s, p->num_path_vertices, key, j->splat, &cn);
                        continue;
                    }
                    if (j->mode == 4) {
                        bdpt_sample(s, p->num_path_vertices, key, px, py, b, j->splat, &cn, NULL, 0, NULL);
                        acc[0] += b[0]; acc[1] += b[1]; acc[2] += b[2];
                        continue;
                    }
                    const int nee = j->mode == 0;
                    eval_contribution(s, p->num_path_vertices, nee ? p->use_mis : 0, nee, key, px, py, b, &ok, &sx, &sy,
                                      sp, &cn);
                    acc[0] += b[0]; acc[1] += b[1]; acc[2] += b[2];
                    if (ok) {
                        double *dst = j->splat + 3 * ((size_t)sy * s->W + sx);
                        atomic_add_double(dst + 0, sp[0]);
                        atomic_add_double(dst + 1, sp[1]);
                        atomic_add_double(dst + 2, sp[2]);
                    }
                }
                double *dst = j->base + 3 * (size_t)pixel;
                dst[0] += acc[0]; dst[1] += acc[1]; dst[2] += acc[2];
            }
    }
    pthread_mutex_lock(&j->mu);
    j->total.ext += cn.ext; j->total.shd += cn.shd;
    pthread_mutex_unlock(&j->mu);
    return NULL;
}
static int render_mode(const orc_scene *s, const orc_render_params *p, int mode, double *base, double *splat,
                       orc_stats *st) {
    if (!s || !p || !base || !splat) return fail("null argument");
    if (mode != 2 && p->num_path_vertices <= 1) { if (st) memset(st, 0, sizeof *st); return 0; } /* pathnextevent.h:163 */
    if (p->spp < 0) return fail("negative spp");
    if ((mode == 0 || mode == 3 || mode == 4) && s->light_cdf.n == 0) return fail("scene has no lights");
    if (mode == 4 && p->num_path_vertices > BDPT_MAXV) return fail("bidirectional path tracing supports at most 64 path vertices");
    if (s->env_on && mode == 4) return fail("the bidirectional path tracer does not support envmap lights");
    if (s->env_on && s->has_medium) return fail("an envmap with a medium is not supported");
    if (s->env_on && mode == 3) return fail("the light tracer does not support envmap lights");
    job_t j;
    memset(&j, 0, sizeof j);
    j.s = s; j.p = p; j.mode = mode; j.base = base; j.splat = splat;
    j.x0 = p->px0; j.y0 = p->py0; j.x1 = p->px1; j.y1 = p->py1;
    if (j.x1 == 0) { j.x0 = 0; j.y0 = 0; j.x1 = s->W; j.y1 = s->H; }
    if (j.x0 < 0 || j.y0 < 0 || j.x1 > s->W || j.y1 > s->H || j.x0 >= j.x1 || j.y0 >= j.y1) return fail("bad pixel window");
    j.tiles_x = (j.x1 - j.x0 + TILE - 1) / TILE;
    j.tiles_y = (j.y1 - j.y0 + TILE - 1) / TILE;
    pthread_mutex_init(&j.mu, NULL);
    int nt = p->threads > 0 ? p->threads : (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (nt < 1) nt = 1;
    if (nt > 256) nt = 256;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    pthread_t th[256];
    for (int i = 0; i < nt; i++) pthread_create(&th[i], NULL, worker, &j);
    for (int i = 0; i < nt; i++) pthread_join(th[i], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    pthread_mutex_destroy(&j.mu);
    if (st) {
        st->paths = (uint64_t)(j.x1 - j.x0) * (uint64_t)(j.y1 - j.y0) * (uint64_t)p->spp;
        st->extend_rays = j.total.ext;
        st->shadow_rays = j.total.shd;
        st->seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    }
    return 0;
}
int orc_render_nee(const orc_scene *s, const orc_render_params *p, double *base, double *splat, orc_stats *st) {
    return render_mode(s, p, 0, base, splat, st);
}
int orc_render_primitive(const orc_scene *s, const orc_render_params *p, double *base, double *splat, orc_stats *st) {
    return render_mode(s, p, 1, base, splat, st);
}
int orc_render_bdpt(const orc_scene *s, const orc_render_params *p, double *base, double *splat, orc_stats *st) {
    return render_mode(s, p, 4, base, splat, st);
}
int orc_bdpt_records(const orc_scene *s, int num_path_vertices, uint32_t seed, int px, int py, uint32_t sample,
                     orc_bdpt_record *out, int max_out, int *n_out) {
    if (!s || !out || !n_out) return fail("null argument");
    if (num_path_vertices < 2 || num_path_vertices > BDPT_MAXV) return fail("num_path_vertices must be in [2, 64]");
    if (s->light_cdf.n == 0 || s->env_on) return fail("needs area lights and no envmap");
    if (px < 0 || py < 0 || px >= s->W || py >= s->H) return fail("pixel outside the film");
    counters cn = {0, 0};
    double b[3];
    *n_out = 0;
    bdpt_sample(s, num_path_vertices, path_key(seed, (uint32_t)(py * s->W + px), sample), px, py, b, NULL, &cn, out,
                max_out, n_out);
    return 0;
}
int orc_render_visualize(const orc_scene *s, const orc_render_params *p, double *base, double *splat, orc_stats *st) {
    return render_mode(s, p, 2, base, splat, st);
}
int orc_render_light(const orc_scene *s, const orc_render_params *p, double *base, double *splat, orc_stats *st) {
    return render_mode(s, p, 3, base, splat, st);
}

/* ------------------------------------------------------------------ trace API */
static void unpack_ray(const float *r, fray *fr) { ray_setup(fr, r, r + 4, r[3], r[7]); }
int orc_trace_closest(const orc_scene *s, const float *rays, uint32_t n, uint32_t *geom, uint32_t *prim, float *t) {
    for (uint32_t i = 0; i < n; i++) {
        fray r; unpack_ray(rays + 8 * (size_t)i, &r);
        double th;
        uint32_t tri = bvh_closest(s, &r, &th);
        if (tri == UINT32_MAX) { geom[i] = UINT32_MAX; prim[i] = UINT32_MAX; t[i] = INFINITY; }
        else { geom[i] = s->tri_geom[tri]; prim[i] = tri - s->geom_first[geom[i]]; t[i] = (float)th; }
    }
    return 0;
}
int orc_trace_any(const orc_scene *s, const float *rays, uint32_t n, uint8_t *occ) {
    for (uint32_t i = 0; i < n; i++) {
        fray r; unpack_ray(rays + 8 * (size_t)i, &r);
        occ[i] = (uint8_t)bvh_any(s, &r);
    }
    return 0;
}
#define NEAR_EPS 1e-5
int orc_trace_closest_brute(const orc_scene *s, const float *rays, uint32_t n, uint32_t *geom, uint32_t *prim,
                            float *t, float *second_t, float *margin) {
    for (uint32_t i = 0; i < n; i++) {
        fray r; unpack_ray(rays + 8 * (size_t)i, &r);
        double best = INFINITY, bm = -1.0; uint32_t bt = UINT32_MAX;
        for (uint32_t k = 0; k < s->n_tris; k++) {
            double th, m, ta;
            if (tri_isect(&r, tri_v(s, k, 0), tri_v(s, k, 1), tri_v(s, k, 2), &th, &m, &ta)) {
                if (th < best) { best = th; bt = k; bm = m; }
            }
        }
        double sec = INFINITY;
        for (uint32_t k = 0; k < s->n_tris; k++) {
            if (k == bt) continue;
            double th, m, ta;
            tri_isect(&r, tri_v(s, k, 0), tri_v(s, k, 1), tri_v(s, k, 2), &th, &m, &ta);
            if (m >= -NEAR_EPS && ta > r.tmin * (1 - NEAR_EPS) && ta <= r.tmax * (1 + NEAR_EPS) && ta < sec) sec = ta;
        }
        if (bt == UINT32_MAX) { geom[i] = UINT32_MAX; prim[i] = UINT32_MAX; t[i] = INFINITY; margin[i] = -1.0f; }
        else { geom[i] = s->tri_geom[bt]; prim[i] = bt - s->geom_first[geom[i]]; t[i] = (float)best; margin[i] = (float)bm; }
        second_t[i] = (float)sec;
    }
    return 0;
}
int orc_trace_any_brute(const orc_scene *s, const float *rays, uint32_t n, uint8_t *occ, uint8_t *amb) {
    for (uint32_t i = 0; i < n; i++) {
        fray r; unpack_ray(rays + 8 * (size_t)i, &r);
        int hit = 0, a = 0;
        for (uint32_t k = 0; k < s->n_tris; k++) {
            double th, m, ta;
            int h = tri_isect(&r, tri_v(s, k, 0), tri_v(s, k, 1), tri_v(s, k, 2), &th, &m, &ta);
            if (h) {
                if (m >= NEAR_EPS && th > r.tmin * (1 + NEAR_EPS) + NEAR_EPS * 1e-3 &&
                    th < r.tmax * (1 - NEAR_EPS)) { hit = 1; }
                else a = 1;
            } else if (m >= -NEAR_EPS && ta > r.tmin * (1 - NEAR_EPS) - 1e-7 && ta <= r.tmax * (1 + NEAR_EPS)) {
                a = 1;
            }
        }
        occ[i] = (uint8_t)(hit || a);
        amb[i] = (uint8_t)(!hit && a);
    }
    return 0;
}
