/*
 * wurst_amd.h -- C-ABI of the MI355X-native `nextevent_path_tracer`.
 *
 * The reference has no FFI: its integrator is the C++ class PathNeeRenderer
 * (src/renderer/pathnextevent.h:15-196) reached from the scene-JSON dispatch
 * `"integrator": "nextevent_path_tracer"` (src/sceneio/simple2parser.h:192-195).
 * This header is the boundary the project's C++ host (and any FFI: ctypes,
 * cgo, JNI) binds; every entry point below names the reference interface it
 * replaces.  Plain pointers and sizes only; nothing throws across the ABI.
 *
 * Conventions
 *   - Every function returns WR_OK (0) or a WR_ERR_* code; wr_last_error()
 *     gives a thread-local message for the last failure on the calling thread.
 *   - Handles are opaque and owned by the caller; destroy them explicitly.
 *   - Host arrays are caller-owned and only read during the call.
 *   - Calls on one scene handle (and films rendered from it) must be
 *     serialised by the caller; distinct scene handles may be used
 *     concurrently (each owns a HIP stream).
 *   - Rays are 8 floats: origin.xyz, tmin, direction.xyz, tmax.  A hit is
 *     accepted for t in (tmin, tmax], the interval rtcIntersect1/rtcOccluded1
 *     use in the reference (src/accel/embreeaccel.h:336-422).
 */
#ifndef WURST_AMD_H
#define WURST_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WR_OK 0
#define WR_ERR_INVALID_ARGUMENT 1
#define WR_ERR_DEVICE 2          /* HIP runtime error */
#define WR_ERR_OUT_OF_MEMORY 3
#define WR_ERR_UNSUPPORTED 4

#define WR_INVALID_ID 0xFFFFFFFFu /* RTC_INVALID_GEOMETRY_ID in the reference */

enum wr_material_type {
    WR_LAMBERT = 0,     /* LambertDiffuseBsdf        src/bsdf/lambertdiffuse.h:7  */
    WR_OREN_NAYAR = 1,  /* OrenNayarBsdf             src/bsdf/orennayar.h:8       */
    WR_PHONG = 2,       /* PhongBsdf (modified)      src/bsdf/phong.h:12          */
    WR_MIXED = 3        /* MixedBsdf{Lambert, Phong} src/bsdf/mixed.h:6           */
};

/* Flattened Bsdf (src/common/bsdf.h:16-43).  twosided != 0 wraps the BSDF in
 * TwoSidedBsdf (src/bsdf/twosided.h:8), as the OBJ loader does for every
 * material (src/sceneio/meshio/assimpio.h:79-121). */
typedef struct {
    int32_t type;
    int32_t twosided;
    float rho_d[3];     /* diffuse reflectance (Lambert, Oren-Nayar, Mixed) */
    float rho_s[3];     /* specular reflectance (Phong, Mixed) */
    float sigma;        /* Oren-Nayar standard deviation */
    float exponent;     /* Phong exponent */
} wr_material;

/* DiffuseAreaLight (src/light/diffusearealight.h:11) over the triangles of
 * geometry `geom`; emits radiance L on the side of Ng = (v1-v0) x (v2-v0). */
typedef struct {
    int32_t geom;
    float L[3];
} wr_light;

/* ThinlensCamera (src/camera/thinlens.h:13-55); aperture_radius 0 = pinhole. */
typedef struct {
    float pos[3], lookat[3], up[3];
    float fovy_deg;
    int32_t width, height;
    float aperture_radius;
    float focal_distance;
} wr_camera;

/* Scene = the triangle meshes the reference attaches to EmbreeAccel
 * (embreeaccel.h:274-315), in attach order: geometry g owns the contiguous
 * triangles with tri_geom == g, and (geomID, primID) = (g, index within g). */
typedef struct {
    const float *verts;            uint32_t n_verts;   /* [n_verts*3] */
    const uint32_t *tris;          uint32_t n_tris;    /* [n_tris*3]  */
    const uint32_t *tri_geom;                          /* [n_tris]    */
    uint32_t n_geoms;
    const int32_t *geom_material;                      /* [n_geoms], -1 for emitters */
    const int32_t *geom_light;                         /* [n_geoms], -1 if not emissive */
    const wr_material *materials;  uint32_t n_materials;
    const wr_light *lights;        uint32_t n_lights;
    wr_camera camera;
} wr_scene_desc;

/* PathNeeRenderer parameters: num_path_vertices / num_spp are the JSON keys of
 * PathNeeRenderer::FromSimple2Json (pathnextevent.h:17-26). */
typedef struct {
    int32_t spp;                 /* samples per pixel rendered by this call */
    int32_t num_path_vertices;   /* N (default 10); max depth = N - 2 */
    uint32_t seed;
    int32_t use_mis;             /* NEE_USE_MIS (pathnextevent.h:13); 1 = on */
    int32_t first_sample;        /* index of the first sample (progressive calls) */
    int32_t max_paths_in_flight; /* 0 = library default (2^27, capped by the samples of a pass) */
    int32_t profile;             /* 1: time each kernel class with HIP events */
    int32_t count_traversal;     /* 1: count visited nodes / tested triangles */
    /* GPU-only scheduling (no reference counterpart; the CPU path has none):
     * 0 = library default; wr_autotune picks them per workload */
    int32_t use_graph;           /* 1: replay the per-iteration launches as a captured hipGraph */
    uint32_t fetch_chunk;        /* traversal work items claimed per atomic (default 256, max 4096) */
    uint32_t refill_lanes;       /* idle lanes that trigger a wave's refill (default 16, max 64) */
} wr_render_params;

typedef struct {
    uint64_t paths;
    uint64_t extend_rays;        /* closest-hit rays traced */
    uint64_t shadow_rays;        /* any-hit rays traced (NEE + light->camera) */
    double wall_ms;              /* host wall time of the call */
    /* with profile = 1: summed HIP-event time and launch count per kernel class.
     * extend_*: the traversal launches, one per pool iteration, each tracing the
     * iteration's closest-hit rays and the previous iteration's shadow rays;
     * shadow_ms and shadow_launches stay 0 (no launch traces shadow rays alone). */
    double extend_ms, shadow_ms, shade_ms, generate_ms, accumulate_ms;
    uint64_t extend_launches, shadow_launches;
    /* with count_traversal = 1 */
    uint64_t extend_nodes, extend_tris, shadow_nodes, shadow_tris;
} wr_render_stats;

typedef struct wr_device wr_device;
typedef struct wr_scene wr_scene;
typedef struct wr_film wr_film;

const char *wr_last_error(void);
/* Library/ABI version string. */
const char *wr_version(void);

/* Binds a GPU (HIP device ordinal).  No reference counterpart (the reference
 * is CPU-only); it stands where Parallel's thread pool is created
 * (src/common/parallel.h:13-21). */
int wr_device_create(int ordinal, wr_device **out);
void wr_device_destroy(wr_device *dev);

/* Checks a scene description without touching a GPU (indices, materials,
 * lights, camera).  The reference reports these as exceptions or asserts
 * (simple2parser.h:90, cdftable.h:36-44 on zero lights, embreeaccel.h:31-61);
 * here they are status codes.  wr_scene_create runs the same checks. */
int wr_scene_validate(const wr_scene_desc *desc);

/* Builds the BVH on the host and verifies its invariants (every triangle in
 * exactly one leaf, child boxes enclose their triangles, stack bound within the
 * traversal stack).  Host only; no GPU needed. */
int wr_bvh_selftest(const wr_scene_desc *desc, uint64_t *n_nodes, uint32_t *max_stack, uint32_t *depth);

/* Uploads geometry, materials, lights and camera and builds the BVH on the GPU.
 * Replaces EmbreeAccel::addGeometry + commit (embreeaccel.h:274-315,466-469),
 * GeneralSubpathSampler's light CdfTable (generalsubpathsampler.h:162-183) and
 * ThinlensCamera construction (thinlens.h:40-55). */
int wr_scene_create(wr_device *dev, const wr_scene_desc *desc, wr_scene **out);
/* Destroy a scene's films and guides first: they refer to it. */
void wr_scene_destroy(wr_scene *scene);
/* BVH builders.  WR_BVH_GPU (default): PLOC clustering (Meister and Bittner
 * 2018) of Morton-ordered triangles on the GPU with SAH leaf selection,
 * collapsed to the 4-wide node layout on the GPU; deterministic.  WR_BVH_HOST:
 * binned SAH on the host, the same layout.  Both stand where rtcCommitScene
 * builds Embree's BVH8 (embreeaccel.h:466-469); any valid BVH gives the same
 * closest hit, so the choice changes build and traversal time only. */
#define WR_BVH_GPU 0
#define WR_BVH_HOST 1
typedef struct {
    int32_t builder;             /* WR_BVH_GPU or WR_BVH_HOST */
    int32_t ploc_radius;         /* GPU builder: clusters searched on each side (0 = default 16, max 64) */
} wr_build_params;

/* wr_scene_create with an explicit builder (params NULL = defaults). */
int wr_scene_create_ex(wr_device *dev, const wr_scene_desc *desc, const wr_build_params *params, wr_scene **out);

typedef struct {
    int32_t builder;
    uint32_t depth;              /* node levels */
    uint32_t max_stack;          /* worst-case traversal stack entries */
    uint32_t iterations;         /* clustering iterations (GPU builder) */
    uint64_t n_nodes;
    double build_ms;             /* wall time of the build, including the mesh upload */
    double sah_cost;             /* expected node + triangle tests of a ray through the root box:
                                  * sum over nodes of area/area(root) + sum over leaves of
                                  * count * area/area(root) */
    uint64_t hash;               /* FNV-1a over the nodes and the leaf-order triangles */
} wr_bvh_stats;

/* Diagnostic (not on the render path): downloads the BVH, checks the invariants
 * of wr_bvh_selftest on it against `desc` (the description the scene was created
 * from) and fills the statistics.  No reference counterpart (Embree's build is
 * internal). */
int wr_scene_bvh_check(wr_scene *scene, const wr_scene_desc *desc, wr_bvh_stats *stats);

/* BVH statistics: nodes, leaf triangles, max stack depth, bytes of nodes+tris. */
int wr_scene_info(const wr_scene *scene, uint64_t *n_nodes, uint64_t *n_tris, uint32_t *max_stack,
                  uint64_t *device_bytes);

/* Homogeneous medium filling the scene (HomogeneousMedium, src/medium/homogeneous.h:11-66;
 * JSON "medium" {"type": "homogeneous", "sigma_a", "sigma_s"}, simple2parser.h:163-167),
 * with a Henyey-Greenstein phase function of asymmetry g in (-1, 1) (g = 0 is the
 * reference's IsotropicPhaseFunction, src/phasefunction/isotropic.h).  Free flights
 * are sampled on one colour channel with the channel-averaged density, as the
 * reference does (homogeneous.h:36-50); NEE shadow rays are weighted by the
 * transmittance.  A null medium (or all-zero coefficients) is vacuum. */
typedef struct {
    float sigma_a[3];
    float sigma_s[3];
    float g;
} wr_medium;
int wr_scene_set_medium(wr_scene *scene, const wr_medium *medium);

/* Envmap light (EnvmapLight over EquiRectRep: light/envmap.h:9-113,
 * sphericalrepresentation/equirectangular.h; JSON "envmap": image path,
 * simple2parser.h:110-119): an equirectangular RGB image rgb[height][width][3],
 * width = 2 height, looked up bilinearly (repeat in u, clamp in v) along
 * u = atan2(-z, -x) / 2pi + 1/2, v = acos(y) / pi (mapping.h:91-96), sampled by
 * a texel CdfTable with weights sin(theta) |rgb|_2, and selected for NEE as the
 * last light with CdfTable weight 1.23 (envmap.h:22-26).  Rays that leave the
 * scene see it (Scene::intersect, scene.h:82-88).  wr_render_nee,
 * wr_render_primitive, wr_render_visualize (a miss counts 1) and batches use it;
 * wr_render_light, wr_render_guided and a medium reject it.  A null rgb removes
 * it.  A scene may have no area light when it has an envmap. */
int wr_scene_set_envmap(wr_scene *scene, const float *rgb, uint32_t width, uint32_t height);

/* FP64 film in device memory: a base buffer (own-pixel samples) and a splat
 * buffer (light->camera connections).  Replaces Film's base and atomic buffers
 * (src/common/film.h:35-68). */
int wr_film_create(wr_scene *scene, wr_film **out);
void wr_film_destroy(wr_film *film);
int wr_film_clear(wr_film *film);
/* Copies the raw sums to host arrays [H][W][3] (row 0 = top of the image).
 * image = base / spp + splat / (spp * W * H)   (pathnextevent.h:154-157). */
int wr_film_read(wr_film *film, double *base, double *splat);

/* Checks render parameters without touching a GPU: spp, first_sample and
 * max_paths_in_flight >= 0, fetch_chunk <= 4096, refill_lanes <= 64.  Every
 * wr_render_* entry point, wr_render_batch and wr_autotune run it before any
 * launch.  The reference takes num_spp / num_path_vertices from the renderer
 * JSON unchecked (pathnextevent.h:17-26); the scheduling fields have no
 * reference counterpart. */
int wr_render_params_validate(const wr_render_params *params);

/* PathNeeRenderer::render (pathnextevent.h:152-191): adds spp samples per pixel
 * of the NEE + MIS estimator into `film`.  Blocking. */
int wr_render_nee(wr_scene *scene, wr_film *film, const wr_render_params *params, wr_render_stats *stats);

/* wr_render_nee without any device allocation.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory during the
 * call (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The film and scene are the caller's; the path pool,
 * per-sample radiance, counters and traversal spill live in a fixed 77 MiB
 * region that the library reserves on each GPU it runs on (held until the
 * process ends), so the call makes no hipMalloc/hipFree, creates no events and
 * instantiates no graph.  Limits: a frame of at most 2^20 pixels (passes of at
 * most 2^20 samples), at most 2^17 paths in flight (max_paths_in_flight 0
 * selects 2^17), profile = 0 and use_graph = 0; anything beyond is rejected
 * with WR_ERR_UNSUPPORTED before any launch.  The film receives what
 * wr_render_nee adds (same samples; the FP64 per-pixel sums are identical
 * when both render the call's spp in one pass, e.g. W * H * spp <= 2^20).
 * All calls on one GPU share the region and leave their contents in it: calls
 * that overlap on one GPU (from different scene streams or threads) race. */
int wr_render_nee_noalloc(wr_scene *scene, wr_film *film, const wr_render_params *params, wr_render_stats *stats);

/* PrimitivePathRenderer::render (primitivepath.h:42-96; JSON "integrator":
 * "primitive_path_tracer", simple2parser.h:184-187): spp samples per pixel of
 * BSDF-sampled camera paths of up to num_path_vertices vertices that add the
 * radiance of every emitter they hit with weight 1 -- no NEE, no light->camera
 * connections (`use_mis` is ignored, the splat buffer stays 0).  The paths
 * scatter in the scene's medium as wr_render_nee's do (the reference's sampler
 * is volumetric here).  Blocking. */
int wr_render_primitive(wr_scene *scene, wr_film *film, const wr_render_params *params, wr_render_stats *stats);

/* LightSplattingRenderer::render (lightsplatting.h:37-113; JSON "integrator":
 * "light_splatting", simple2parser.h:188-191): spp * W * H light subpaths of up
 * to num_path_vertices - 1 vertices (light vertex, then cosine emission and BSDF
 * sampling; an emitter hit ends one); the light vertex and every surface vertex
 * connect to the camera with weight 1 and add G We1 f T to the splat buffer at
 * the camera raster: image = splat / (spp * W * H), the base buffer stays 0.
 * `use_mis` is ignored; the medium is ignored (the reference's sampler here is
 * not volumetric).  Blocking. */
int wr_render_light(wr_scene *scene, wr_film *film, const wr_render_params *params, wr_render_stats *stats);

/* VisualizeGeometryNormal::render (visualizegeomnormal.h:32-74; JSON
 * "integrator": "visualize", simple2parser.h:196-199): per sample the camera
 * ray's first hit adds (1, 1, 1) on an emitter and the unit geometric normal
 * (v1 - v0) x (v2 - v0), unflipped, as RGB on any other surface; misses add 0;
 * image = base / spp.  Only spp, seed, first_sample and profile are read; the
 * medium is ignored (the reference's sampler is not volumetric here).
 * Blocking. */
int wr_render_visualize(wr_scene *scene, wr_film *film, const wr_render_params *params, wr_render_stats *stats);

/* BidirPathRenderer::render (bidirpath.h:316-395; JSON "integrator":
 * "bidirectional_path_tracer", simple2parser.h:176-179): per camera sample one
 * light subpath and one camera subpath (GeneralSubpathSampler, not volumetric)
 * and every strategy (s, t) with s + t <= num_path_vertices, weighted by the
 * balance heuristic (computeMisWeight, power 1).  t = 1 connections go to the
 * splat buffer at their raster, the rest to the base buffer at the pixel:
 * image = base / spp + splat / (spp * W * H), the NEE film layout.  N <= 1
 * renders nothing; N above 64, an envmap or a scene without lights is an
 * error; `use_mis` is ignored (the reference always weights) and so is the
 * medium.  The light subpath stops at N - 1 vertices (an N-th, which the
 * reference traces, enters no strategy).  Keeps a per-scene buffer of
 * 64 B x (N - 1) light-vertex records per traversal lane.  Blocking. */
int wr_render_bdpt(wr_scene *scene, wr_film *film, const wr_render_params *params, wr_render_stats *stats);

/* Closest hit for n host rays (EmbreeAccel::intersect -> rtcIntersect1,
 * embreeaccel.h:353-422).  Misses give geom = prim = WR_INVALID_ID, t = +inf. */
int wr_trace_closest(wr_scene *scene, const float *rays, uint32_t n, uint32_t *geom, uint32_t *prim, float *t);
/* Any hit (EmbreeAccel::isIntersect -> rtcOccluded1, embreeaccel.h:336-351). */
int wr_trace_any(wr_scene *scene, const float *rays, uint32_t n, uint8_t *occluded);

/* Device-resident variants (pointers from hipMalloc, e.g. torch tensors):
 * rays [n*8] float, hits [n] of {uint32 leaf-tri or WR_INVALID_ID, float t}.
 * Launched on the scene's stream; `ms` (nullable) receives the kernel time
 * measured with HIP events on that stream. */
int wr_trace_closest_device(wr_scene *scene, const float *d_rays, uint32_t n, void *d_hits, float *ms);
int wr_trace_any_device(wr_scene *scene, const float *d_rays, uint32_t n, uint8_t *d_occluded, float *ms);

/* ------------------------------------------------------------------ per-call checks
 * The device functions the kernels inline, evaluated one record per GPU lane on
 * host inputs (staged through device buffers allocated per call), so they can
 * be compared call by call with the reference's classes.  Diagnostics; not on
 * the render path.  Vectors are float[3], uniforms float[2], n records each. */
/* Bsdf::evalBsdf / evalPdfW (bsdf.h:20-42; lambertdiffuse.h:16-51,
 * orennayar.h:14-83, phong.h:22-93, mixed.h:20-60, twosided.h:25-63) in the
 * local frame (y = normal): f [n*3] and pdf [n] for (wi, wo), material per record. */
int wr_debug_bsdf_eval(wr_device *dev, const wr_material *mats, const float *wi, const float *wo, uint32_t n, float *f,
                       float *pdf);
/* Bsdf::sampleBsdfCosByPdf: sampled direction wo [n*3], weight f cos / pdf
 * [n*3] and pdf [n] for incoming wi and uniforms u. */
int wr_debug_bsdf_sample(wr_device *dev, const wr_material *mats, const float *wi, const float *u, uint32_t n, float *wo,
                         float *weight, float *pdf);
/* createLightSubpath with one vertex (generalsubpathsampler.h:199-255 ->
 * DiffuseAreaLight::sampleLe0ByPdf, diffusearealight.h:53-72): point, unit
 * normal, contribution (pi L / pdfA_point) / pdfLight and pdfA for uniforms u
 * on the scene's area lights. */
int wr_debug_light_sample(wr_scene *scene, const float *u, uint32_t n, float *pos, float *nrm, float *contrib,
                          float *pdfA);
/* ThinlensCamera::sampleWe0ByPdf + samplePerPixelWe1CosByPdf (thinlens.h:121-166):
 * lens point org, unit direction dir and per-pixel pdfW of the sample at
 * pixel (x, y) = pixel[2i..2i+1] with lens uniforms ulens and jitter upix. */
int wr_debug_camera_ray(wr_scene *scene, const uint32_t *pixel, const float *ulens, const float *upix, uint32_t n,
                        float *org, float *dir, float *pdfW);
/* ThinlensCamera::evalWe1 (thinlens.h:62-80) from lens point org (world) toward
 * direction w: We1 [n] (0 outside the film) and the raster position [n*2]. */
int wr_debug_camera_we1(wr_scene *scene, const float *org, const float *w, uint32_t n, float *we1, float *raster);
/* Measured ceiling of the rate at which a CU's L1 serves scattered lane
 * requests, the bound of the traversal kernels (each lane fetches its own BVH
 * node in 16-byte rows).  No counterpart in the reference: a hardware
 * diagnostic for the benchmark's roofline.  A full grid (every CU at full
 * occupancy) in which each lane issues `loads_per_lane` independent 16-byte
 * loads, eight in flight, at pseudo-random 64-byte-aligned offsets of a
 * device buffer of `footprint` bytes (a power of two, 64 B .. 1 GiB; a few MB
 * stay L2-resident).  Reports the timed launch's duration (ms), the loads it
 * issued and the GPU's compute units. */
int wr_debug_l1_request_probe(wr_device *dev, uint64_t footprint, uint32_t loads_per_lane, double *ms, uint64_t *loads,
                              uint32_t *n_cu);

/* ------------------------------------------------------------------ batches
 * Several independent renders (scene, film, spp, depth, seed) in one path pool:
 * every iteration is one launch of each wavefront kernel for all jobs, which
 * keeps the GPU full when single jobs are too small to.  No reference
 * counterpart (the reference renders one scene per process).  Each job's film
 * receives exactly what wr_render_nee with the same parameters would add
 * (same random numbers per pixel and sample; when a job has an envmap, to FP32
 * rounding).  Scenes with a medium are rejected.  All scenes must live on the
 * same device; the batch runs on the first job's scene stream and path pool,
 * so calls touching those scenes must be serialised with it.  From `params`
 * only max_paths_in_flight, profile, use_graph, fetch_chunk and refill_lanes
 * are used. */
typedef struct {
    wr_scene *scene;
    wr_film *film;               /* a film of `scene` */
    int32_t spp;
    int32_t num_path_vertices;   /* 2..255 */
    uint32_t seed;
    int32_t use_mis;
    int32_t first_sample;
} wr_batch_job;
int wr_render_batch(const wr_batch_job *jobs, uint32_t n_jobs, const wr_render_params *params, wr_render_stats *stats);

/* Picks the path-pool size, traversal work chunk, refill threshold and graph
 * replay for a batch by timing short trial renders (1 spp per job, into
 * scratch films) of each candidate; writes the fastest into *best (other
 * fields copied from *base) and its rays per second into *mrays. */
int wr_autotune(const wr_batch_job *jobs, uint32_t n_jobs, const wr_render_params *base, wr_render_params *best,
                double *mrays);

/* ------------------------------------------------------------------ guiding
 * Learned path guiding.  No reference counterpart: the reference has no
 * guiding; PathNeeRenderer samples the BSDF only (generalsubpathsampler.h:81-130).
 * The network (one-blob 32 -> 4 x 64 ReLU -> 40) outputs 8 normalised
 * spherical-Gaussian (von Mises-Fisher) lobes: for lobe k the raw outputs
 * 5k..5k+4 are (weight logit, axis xyz, sharpness), with weights softmax(logits),
 * axes normalised and sharpness softplus(raw) clamped to [0.05, 2000].  Its
 * density, lambda / (2 pi (1 - e^{-2 lambda})) exp(lambda (mu.w - 1)) per lobe, is
 * mixed with BSDF sampling by one-sample MIS, so the estimator stays unbiased
 * for any weights.  Parameters are FP32 (layer l: W_l [in][out] row-major, then
 * b_l); the GPU trains in BF16 with FP32 accumulation and infers in BF16 or,
 * optionally, e4m3 (FP8). */
#define WR_GUIDE_NUM_PARAMS 17192
#define WR_GUIDE_NUM_OUTPUTS 40

typedef struct wr_guide wr_guide;

typedef struct {
    float learning_rate;       /* Adam step size (1e-2) */
    float beta1, beta2, epsilon; /* Adam moments (0.9, 0.999, 1e-8) */
    uint32_t batch_size;       /* records per training step (2^18) */
    uint32_t record_capacity;  /* training records kept per pass (2^20) */
    uint32_t seed;             /* weight initialisation */
    float avg_decay;           /* the renderer's inference uses the exponential average of the
                                  parameters over the Adam steps, avg = decay avg + (1 - decay)
                                  params after each step, zero-started at creation and
                                  wr_guide_set_params and divided by 1 - decay^steps (0.99);
                                  0: inference uses the parameters themselves */
} wr_guide_config;

typedef struct {
    float guide_fraction;      /* probability of sampling the network at a vertex (0.5) */
    uint32_t spp_per_pass;     /* samples per pixel between training rounds */
    uint32_t train_steps;      /* Adam steps after each pass (0: render only) */
    uint32_t warmup_passes;    /* passes that sample the BSDF only (and collect); the network
                                  is sampled from pass warmup_passes on, trained or not */
    int32_t fp8_inference;     /* 1: the renderer's network inference runs the e4m3 variant
                                  (wr_guide_forward_fp8); training stays BF16 */
    uint32_t max_guided_bounce; /* vertices at bounce < this (0 = the camera ray's hit) are guided
                                   and recorded; deeper ones sample the BSDF or phase function
                                   only (probability 0 of the network, so still unbiased).
                                   0 = every bounce */
    uint32_t overlap_training; /* 1: pass q renders with the network trained on pass q - 2, so
                                  that the training on pass q - 1 runs beside it (give such a
                                  render two warm-up passes); 0: pass q waits for that training */
} wr_guided_params;

typedef struct {
    uint64_t records;          /* training records collected */
    uint64_t train_steps;      /* Adam steps taken */
    double train_ms;           /* HIP-event time of finalise + training */
    double last_loss;          /* normalised loss of the last step */
} wr_guide_stats;

void wr_guide_config_default(wr_guide_config *cfg);
/* A network bound to a scene (its bounds fix the position encoding) and that
 * scene's stream. */
int wr_guide_create(wr_scene *scene, const wr_guide_config *cfg, wr_guide **out);
void wr_guide_destroy(wr_guide *guide);
/* Host copies of the FP32 parameters; set also resets the Adam state. */
int wr_guide_get_params(wr_guide *guide, float *params);
int wr_guide_set_params(wr_guide *guide, const float *params);
/* Network outputs for n host records (pos, nrm, wo: n*3 floats; nrm faces the
 * incoming ray, wo points to the previous vertex): out n*40 raw outputs. */
int wr_guide_forward(wr_guide *guide, const float *pos, const float *nrm, const float *wo, uint32_t n, float *out);
/* The same through the FP8 inference variant: e4m3 weights with a power-of-two
 * scale per layer (largest |weight| scaled into (64, 128]) and e4m3 activations
 * with a power-of-two scale per layer and record (its largest input to the
 * layer likewise), FP32 accumulation and FP32 biases. */
int wr_guide_forward_fp8(wr_guide *guide, const float *pos, const float *nrm, const float *wo, uint32_t n, float *out);
/* Gradient of the loss (1/n) sum_r -weight_r log q_r(wi_r) over n host records
 * (wi: n*3 sampled directions, weight: n); grad has WR_GUIDE_NUM_PARAMS floats. */
int wr_guide_gradient(wr_guide *guide, const float *pos, const float *nrm, const float *wo, const float *wi,
                      const float *weight, uint32_t n, float *grad, double *loss);
/* `steps` Adam steps on n host records (the whole set per step). */
int wr_guide_train(wr_guide *guide, const float *pos, const float *nrm, const float *wo, const float *wi,
                   const float *weight, uint32_t n, uint32_t steps, double *loss);
/* The mixture of n raw output rows (n*40): pdf at dirs (n*3), and one sample
 * per row from u (n*3 uniforms) with its pdf. */
int wr_guide_mixture(wr_guide *guide, const float *raw, const float *dirs, const float *u, uint32_t n, float *pdf_dirs,
                     float *samples, float *sample_pdf);
/* wr_render_nee with guided sampling: progressive passes of spp_per_pass;
 * after each pass the network trains on records of that pass. */
int wr_render_guided(wr_scene *scene, wr_film *film, wr_guide *guide, const wr_render_params *params,
                     const wr_guided_params *gparams, wr_render_stats *stats, wr_guide_stats *gstats);

#ifdef __cplusplus
}
#endif
#endif
