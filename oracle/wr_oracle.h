/*
 * wr_oracle.h -- CPU FP64 restatement of wurstrenderer's `nextevent_path_tracer`.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or
 * executed only by tests/, by __graft_entry__.smoke() and by bench.py's
 * `cpu_baseline` leg.  The product path (wurstrenderer_amd/) never calls it
 * and never falls back to it.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the reference checkout).  All arithmetic is FP64, like the reference
 * (README.md:83); geometry enters as FP32 vertices, rays are rounded to FP32
 * before intersection exactly as EmbreeAccel does (src/accel/embreeaccel.h:338-347,
 * 356-366), and the intersection itself is evaluated in FP64.
 *
 * Pinning (what checks each function against something other than itself):
 *   rng            : uniformity / independence statistics (tests/test_oracle_math.py)
 *   onb, mappings  : reference golden vectors (tests/golden/ref_vectors.json),
 *                    orthonormality, closed-form pdf integrals
 *   bsdf_*         : reference golden vectors + Monte-Carlo pdf normalisation +
 *                    reciprocity of eval
 *   light_*, cdf   : reference golden vectors + area uniformity
 *   camera_*       : reference golden vectors + raster round trip
 *   trace_*        : brute force over all triangles; closed-form hits
 *   render_nee     : MIS on/off agreement (pins MIS weights), closed-form
 *                    direct illumination of a plane by a polygonal luminaire
 *                    (Lambert's formula), directly-visible emitter radiance,
 *                    furnace-style energy bound.
 *   medium, HG     : HG normalisation and mean cosine g (quadrature / Monte
 *                    Carlo), free-flight mean 1/sigma_t; renders: a purely
 *                    absorbing medium seen against an emitting enclosure gives
 *                    L exp(-sigma_t d) per pixel, a purely scattering one in an
 *                    emitting enclosure gives L (radiative equilibrium), MIS
 *                    on/off agreement in a scattering medium.
 * The full integrator of the reference cannot be built here (its accelerator,
 * Embree 3.5.2, needs CMake-generated headers), so there are no reference film
 * fixtures; render_nee is pinned by the closed forms above.
 */
#ifndef WR_ORACLE_H
#define WR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_LAMBERT = 0, ORC_OREN_NAYAR = 1, ORC_PHONG = 2, ORC_MIXED = 3 };

typedef struct {
    int32_t type;        /* ORC_* */
    int32_t twosided;    /* wrap in TwoSidedBsdf (src/bsdf/twosided.h) */
    float rho_d[3];      /* Lambert / Oren-Nayar diffuse reflectance */
    float rho_s[3];      /* Phong specular reflectance */
    float sigma;         /* Oren-Nayar std deviation (raw, as orennayar.h:52) */
    float exponent;      /* Phong exponent */
} orc_material;          /* 40 bytes */

typedef struct {
    int32_t geom;        /* emissive geometry (its triangles are the emitter) */
    float L[3];          /* radiance */
} orc_light;             /* 16 bytes */

typedef struct {
    float pos[3], lookat[3], up[3];
    float fovy_deg;
    int32_t width, height;
    float aperture_radius;   /* 0 = pinhole */
    float focal_distance;
} orc_camera;                /* 56 bytes */

typedef struct {
    const float *verts;        uint32_t n_verts;   /* [n_verts*3] */
    const uint32_t *tris;      uint32_t n_tris;    /* [n_tris*3]  */
    const uint32_t *tri_geom;                      /* [n_tris] geometry id (contiguous per geometry) */
    uint32_t n_geoms;
    const int32_t *geom_material;                  /* [n_geoms] material index, -1 for emitters */
    const int32_t *geom_light;                     /* [n_geoms] light index, -1 if not emissive */
    const orc_material *materials; uint32_t n_materials;
    const orc_light *lights;       uint32_t n_lights;
    orc_camera camera;
} orc_scene_desc;

typedef struct orc_scene orc_scene;

typedef struct {
    int32_t spp;                 /* samples per pixel in this call */
    int32_t num_path_vertices;   /* N, pathnextevent.h:23 (default 10) */
    uint32_t seed;
    int32_t use_mis;             /* NEE_USE_MIS (pathnextevent.h:13) */
    int32_t first_sample;        /* sample index of the first spp (progressive rendering) */
    int32_t threads;             /* 0 = all hardware threads */
    int32_t px0, py0, px1, py1;  /* pixel window; px1 == 0 means the whole film */
} orc_render_params;

typedef struct {
    uint64_t paths;
    uint64_t extend_rays;
    uint64_t shadow_rays;
    double seconds;
} orc_stats;

int  orc_scene_create(const orc_scene_desc *desc, orc_scene **out);
void orc_scene_destroy(orc_scene *s);
const char *orc_last_error(void);

/* Adds the estimator's sums into film_base / film_splat ([H][W][3] doubles each).
 * image = base / spp_total + splat / (spp_total * W * H)   (pathnextevent.h:154-157) */
int orc_render_nee(const orc_scene *s, const orc_render_params *p,
                   double *film_base, double *film_splat, orc_stats *stats);
/* PrimitivePathRenderer (primitivepath.h:42-96): the same walk without NEE or
 * light->camera connections, emitter hits at weight 1 (use_mis is ignored;
 * film_splat is left unchanged). */
int orc_render_primitive(const orc_scene *s, const orc_render_params *p,
                         double *film_base, double *film_splat, orc_stats *stats);
/* VisualizeGeometryNormal (visualizegeomnormal.h:32-74): per sample the primary
 * hit's unit geometric normal, or (1, 1, 1) on an emitter; image = base / spp.
 * num_path_vertices and use_mis are ignored; the medium is not traversed. */
int orc_render_visualize(const orc_scene *s, const orc_render_params *p,
                         double *film_base, double *film_splat, orc_stats *stats);
/* LightSplattingRenderer (lightsplatting.h:37-113): spp * W * H light subpaths
 * connected to the camera; image = splat / (spp * W * H).  use_mis and the
 * medium are ignored (the reference's sampler there is not volumetric). */
int orc_render_light(const orc_scene *s, const orc_render_params *p,
                     double *film_base, double *film_splat, orc_stats *stats);
/* BidirPathRenderer (bidirpath.h:44-421): per camera sample one light and one
 * camera subpath and every strategy (s, t) with s + t <= N, balance-heuristic
 * weights; t = 1 into film_splat, the rest into film_base (the NEE film
 * layout).  N <= 64; use_mis and the medium are ignored (the reference's BDPT
 * always weights, and its subpath sampler is not volumetric); no envmap. */
int orc_render_bdpt(const orc_scene *s, const orc_render_params *p,
                    double *film_base, double *film_splat, orc_stats *stats);
/* Test hook: the strategies of one camera sample of orc_render_bdpt with a
 * nonzero unweighted contribution, each with its connected path (vertex 1 the
 * light point ... vertex n the camera; kind 0 camera, 1 light, 2 surface; id
 * the material or light index), the area pdfs lp / cp with which the light /
 * camera side makes each vertex as the weight uses them, and the weight. */
typedef struct {
    int s, t, n, pad;
    double weight, contrib[3];
    double pos[64][3], nrm[64][3];
    int kind[64], id[64];
    double lp[64], cp[64];
} orc_bdpt_record;
int orc_bdpt_records(const orc_scene *s, int num_path_vertices, uint32_t seed, int px, int py, uint32_t sample,
                     orc_bdpt_record *out, int max_out, int *n_out);

/* A homogeneous medium filling the whole scene (src/medium/homogeneous.h:11-66,
 * JSON "medium" {"sigma_a", "sigma_s"}, simple2parser.h:163-167) with a
 * Henyey-Greenstein phase function of asymmetry g (0 = isotropic,
 * src/phasefunction/isotropic.h).  All-zero coefficients remove it. */
int orc_scene_set_medium(orc_scene *s, const double sigma_a[3], const double sigma_s[3], double g);
/* Envmap light (EnvmapLight over EquiRectRep, light/envmap.h,
 * sphericalrepresentation/equirectangular.h): rgb[h][w][3] FP32, w = 2 h; a null
 * rgb removes it.  The last entry of the light CdfTable, weight 1.23. */
int orc_scene_set_envmap(orc_scene *s, const float *rgb, int w, int h);
/* EquiRectRep::eval (bilinear) and ::pdfA at direction d. */
void orc_env_eval(const orc_scene *s, const double d[3], double rgb[3], double *pdf);
/* EquiRectRep::sample with the reference's use of u (the in-texel u is the
 * remapped texel-selection number): direction, solid-angle pdf, eval / pdf. */
void orc_env_sample_ref(const orc_scene *s, const double u[2], double dir[3], double *pdf, double value[3]);
double orc_hg_pdf(double g, double cos_theta);
void orc_hg_sample(double g, const double fwd[3], const double u[2], double out[3]);

/* Rays are 8 floats: o.xyz, tmin, d.xyz, tmax.  Hit interval is (tmin, tmax]. */
int orc_trace_closest(const orc_scene *s, const float *rays, uint32_t n,
                      uint32_t *geom, uint32_t *prim, float *t);
int orc_trace_any(const orc_scene *s, const float *rays, uint32_t n, uint8_t *occluded);
/* Brute force over every triangle; also reports tie information:
 * second_t  = smallest t of any other triangle that hits or nearly hits
 *             (min barycentric >= -1e-5), +inf if none;
 * margin    = min barycentric of the closest hit (-1 on miss). */
int orc_trace_closest_brute(const orc_scene *s, const float *rays, uint32_t n,
                            uint32_t *geom, uint32_t *prim, float *t,
                            float *second_t, float *margin);
int orc_trace_any_brute(const orc_scene *s, const float *rays, uint32_t n,
                        uint8_t *occluded, uint8_t *ambiguous);

/* ---- per-call functions (golden-vector interface), all FP64, local frame y = normal ---- */
uint32_t orc_rng_u32(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim);
double   orc_rng_uniform(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim);
void orc_onb(const double n[3], double x[3], double z[3]);
void orc_cosine_hemisphere(const double u[2], double out[3]);
void orc_barycentric_from_square(const double u[2], double b[2]);
void orc_disk_from_square(const double u[2], double d[2]);
/* CdfTable::sample over weights; returns index, writes pdf and remapped sample */
int  orc_cdf_sample(const double *weights, int n, double u, double *pdf, double *remapped);
void orc_bsdf_eval(const orc_material *m, const double wi[3], const double wo[3],
                   double f[3], double *pdf);
void orc_bsdf_eval_n(const orc_material *m, int n, const double *wi, const double *wo, double *f, double *pdf);
/* returns f*cos/pdf in w[3]; the pdf written is the one the reference reports */
void orc_bsdf_sample(const orc_material *m, const double wi[3], const double u[2],
                     double wo[3], double w[3], double *pdf);
/* light sampling: u = the two uniforms of createLightSubpath */
int  orc_light_sample(const orc_scene *s, const double u[2], double pos[3], double normal[3],
                      double contrib[3], double *pdfA);
/* camera: per-pixel primary ray and film-wide pdfW; evalWe1 raster projection */
void orc_camera_ray(const orc_scene *s, int px, int py, const double ulens[2], const double upix[2],
                    double org[3], double dir[3], double raster[2], double *pdfW);
double orc_camera_we1(const orc_scene *s, const double cam_pos[3], const double w[3], double raster[2]);

#ifdef __cplusplus
}
#endif
#endif
