// wurst_host.cpp -- scene JSON, OBJ/MTL loading and the GPU PathNeeRenderer.
#include "wurst_host.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

namespace wurst {

// ------------------------------------------------------------------ film
std::vector<float> Film::image() const {
    std::vector<float> img(base.size());
    for (size_t i = 0; i < base.size(); i++) img[i] = float(base[i] * base_scale + splat[i] * splat_scale);
    return img;
}

void Film::save_pfm(const std::string &path) const {
    std::ofstream os(path, std::ios::binary);
    if (!os.is_open()) throw std::runtime_error("cannot open file " + path);
    os << "PF\n" << W << " " << H << "\n-1\n";
    const std::vector<float> img = image();
    for (int y = H - 1; y >= 0; y--) os.write(reinterpret_cast<const char *>(&img[3 * size_t(y) * W]), sizeof(float) * 3 * W);
}

// ------------------------------------------------------------------ camera
// ThinlensCamera::FromSimple2ParseJson (thinlens.h:13-26).  The reference reads
// the key "radius" for both the aperture radius (default 0) and the focal
// distance (default 1); that reading is kept, so "radius": 0 is rejected
// (focal distance 0) instead of producing NaN directions.
std::shared_ptr<ThinlensCamera> ThinlensCamera::FromSimple2ParseJson(const wj::Value &j) {
    auto cam = std::make_shared<ThinlensCamera>();
    const auto pos = j.vec3("pos"), at = j.vec3("lookat"), up = j.vec3("up");
    const double fovy = j.number("fovy", 90.0);
    if (fovy > 180.0) throw std::runtime_error("Thinlens camera doesn't support when fov > Pi radians");
    const auto res = j.ivec2("resolution");
    const double aperture = j.number("radius", 0.0);
    const double focal = j.number("radius", 1.0);
    if (!(focal > 0.0)) throw std::runtime_error("camera: focal distance (key \"radius\") must be > 0");
    if (res[0] <= 0 || res[1] <= 0) throw std::runtime_error("camera: bad resolution");
    wr_camera &d = cam->desc;
    for (int k = 0; k < 3; k++) { d.pos[k] = float(pos[k]); d.lookat[k] = float(at[k]); d.up[k] = float(up[k]); }
    d.fovy_deg = float(fovy);
    d.width = res[0];
    d.height = res[1];
    d.aperture_radius = float(aperture);
    d.focal_distance = float(focal);
    cam->film = std::make_shared<Film>(res[0], res[1]);
    return cam;
}

wr_scene_desc SceneData::desc(const wr_camera &camera) const {
    wr_scene_desc d{};
    d.verts = verts.data();
    d.n_verts = uint32_t(verts.size() / 3);
    d.tris = tris.data();
    d.n_tris = uint32_t(tri_geom.size());
    d.tri_geom = tri_geom.data();
    d.n_geoms = uint32_t(geom_material.size());
    d.geom_material = geom_material.data();
    d.geom_light = geom_light.data();
    d.materials = materials.data();
    d.n_materials = uint32_t(materials.size());
    d.lights = lights.data();
    d.n_lights = uint32_t(lights.size());
    d.camera = camera;
    return d;
}

// ------------------------------------------------------------------ OBJ / MTL
namespace {

struct MtlEntry {
    double Kd[3] = {0, 0, 0}, Ks[3] = {0, 0, 0}, Ke[3] = {0, 0, 0};
    double Ns = 0.0;
};

// SpectrumConvertUtil::FromSrgb (src/common/math/spectrum.h:13-17): the sRGB transfer curve
double from_srgb(double v) { return v <= 0.04045 ? v * (1.0 / 12.92) : std::pow((v + 0.055) * (1.0 / 1.055), 2.4); }

std::string dir_of(const std::string &p) {
    const size_t s = p.find_last_of('/');
    return s == std::string::npos ? std::string(".") : p.substr(0, s);
}

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\r')) a++;
    while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r')) b--;
    return s.substr(a, b - a);
}

void load_mtl(const std::string &path, std::map<std::string, MtlEntry> *out) {
    std::ifstream is(path);
    if (!is.is_open()) throw std::runtime_error("cannot open material library " + path);
    MtlEntry *cur = nullptr;
    for (std::string line; std::getline(is, line);) {
        std::istringstream ss(line);
        std::string key;
        ss >> key;
        if (key == "newmtl") {
            std::string name;
            std::getline(ss, name);
            cur = &(*out)[trim(name)];
        } else if (cur && (key == "Kd" || key == "Ks" || key == "Ke")) {
            double *dst = key == "Kd" ? cur->Kd : (key == "Ks" ? cur->Ks : cur->Ke);
            ss >> dst[0];
            if (!(ss >> dst[1] >> dst[2])) dst[1] = dst[2] = dst[0];
        } else if (cur && key == "Ns") {
            ss >> cur->Ns;
        }
    }
}

int parse_index(const std::string &tok, int n_verts) {
    const int i = std::atoi(tok.c_str());  // "v", "v/t", "v//n", "v/t/n": the leading integer
    if (i > 0) return i - 1;
    if (i < 0) return n_verts + i;
    throw std::runtime_error("obj: bad face index " + tok);
}

}  // namespace

void LoadObj(const std::string &path, bool srgb, SceneData *scene) {
    std::ifstream is(path);
    if (!is.is_open()) throw std::runtime_error("cannot open file " + path);
    std::vector<double> pos;
    std::map<std::string, MtlEntry> mtl;
    struct Group {
        std::string material;
        std::vector<uint32_t> faces;  // global vertex indices, 3 per triangle
    };
    std::vector<Group> groups;
    std::string cur_mat, cur_obj;
    bool new_group = true;
    for (std::string line; std::getline(is, line);) {
        std::istringstream ss(line);
        std::string key;
        ss >> key;
        if (key == "v") {
            double x, y, z;
            ss >> x >> y >> z;
            pos.push_back(x); pos.push_back(y); pos.push_back(z);
        } else if (key == "f") {
            std::vector<int> idx;
            for (std::string tok; ss >> tok;) idx.push_back(parse_index(tok, int(pos.size() / 3)));
            if (idx.size() < 3) throw std::runtime_error("obj: face with fewer than 3 vertices");
            if (new_group) { groups.push_back({cur_mat, {}}); new_group = false; }
            for (size_t k = 1; k + 1 < idx.size(); k++) {  // fan triangulation
                for (int v : {idx[0], idx[k], idx[k + 1]}) {
                    if (v < 0 || size_t(v) >= pos.size() / 3) throw std::runtime_error("obj: face index out of range");
                    groups.back().faces.push_back(uint32_t(v));
                }
            }
        } else if (key == "usemtl") {
            std::string m;
            std::getline(ss, m);
            m = trim(m);
            if (m != cur_mat) { cur_mat = m; new_group = true; }
        } else if (key == "o" || key == "g") {
            std::string o;
            std::getline(ss, o);
            if (trim(o) != cur_obj) { cur_obj = trim(o); new_group = true; }
        } else if (key == "mtllib") {
            std::string f;
            std::getline(ss, f);
            load_mtl(dir_of(path) + "/" + trim(f), &mtl);
        }
    }
    // materials -> BSDFs / emitters (assimpio.h:79-121).  Geometry ids follow
    // the reference's attach order: the OBJ's surfaces first, then its emitters
    // (AssimpInOut::LoadFromObj splits them, assimpio.h:123-150, and
    // Simple2Parser::ParseScene attaches mGeometries before mLightGeometries,
    // simple2parser.h:38-61), each in file order.
    auto is_emissive = [&](const Group &g) {
        auto it = mtl.find(g.material);
        if (it == mtl.end()) throw std::runtime_error("obj: unknown material \"" + g.material + "\"");
        const MtlEntry &m = it->second;
        return m.Ke[0] != 0.0 || m.Ke[1] != 0.0 || m.Ke[2] != 0.0;
    };
    std::vector<const Group *> order;
    for (int pass = 0; pass < 2; pass++)
        for (const Group &g : groups)
            if (!g.faces.empty() && is_emissive(g) == (pass == 1)) order.push_back(&g);
    for (const Group *gp : order) {
        const Group &g = *gp;
        const MtlEntry &m = mtl.find(g.material)->second;
        const bool emissive = is_emissive(g);
        const uint32_t geom = uint32_t(scene->geom_material.size());
        if (emissive) {
            if (m.Ns != 0.0) throw std::runtime_error("obj: emitter with a Phong exponent (cone-emission area light) is not supported");
            wr_light l{};
            l.geom = int32_t(geom);
            for (int k = 0; k < 3; k++) l.L[k] = float(m.Ke[k]);
            scene->lights.push_back(l);
            scene->geom_material.push_back(-1);
            scene->geom_light.push_back(int32_t(scene->lights.size() - 1));
        } else {
            wr_material wm{};
            double kd[3], ks[3];
            for (int k = 0; k < 3; k++) {
                kd[k] = srgb ? from_srgb(m.Kd[k]) : m.Kd[k];
                ks[k] = srgb ? from_srgb(m.Ks[k]) : m.Ks[k];
            }
            const bool has_d = kd[0] != 0.0 || kd[1] != 0.0 || kd[2] != 0.0;
            const bool has_s = ks[0] != 0.0 || ks[1] != 0.0 || ks[2] != 0.0;
            wm.type = !has_d ? WR_PHONG : (!has_s ? WR_LAMBERT : WR_MIXED);
            wm.twosided = 1;
            for (int k = 0; k < 3; k++) { wm.rho_d[k] = float(kd[k]); wm.rho_s[k] = float(ks[k]); }
            wm.exponent = float(m.Ns / 4.0);  // Assimp reports Ns x 4; the reference divides it back (assimpio.h:310-315)
            scene->materials.push_back(wm);
            scene->geom_material.push_back(int32_t(scene->materials.size() - 1));
            scene->geom_light.push_back(-1);
        }
        // per-geometry vertex list, in first-use order
        std::map<uint32_t, uint32_t> remap;
        const uint32_t vbase = uint32_t(scene->verts.size() / 3);
        for (uint32_t v : g.faces) {
            auto r = remap.find(v);
            uint32_t local;
            if (r == remap.end()) {
                local = uint32_t(remap.size());
                remap[v] = local;
                for (int k = 0; k < 3; k++) scene->verts.push_back(float(pos[3 * size_t(v) + k]));
            } else {
                local = r->second;
            }
            scene->tris.push_back(vbase + local);
        }
        for (size_t t = 0; t < g.faces.size() / 3; t++) scene->tri_geom.push_back(geom);
    }
}

// ------------------------------------------------------------------ rendering
RenderContext::~RenderContext() {
    if (device) wr_device_destroy(device);
}

wr_device *RenderContext::get() {
    if (!device && wr_device_create(device_ordinal, &device) != WR_OK) throw std::runtime_error(wr_last_error());
    return device;
}

static void check(int rc) {
    if (rc != WR_OK) throw std::runtime_error(wr_last_error());
}

std::shared_ptr<PathNeeRenderer> PathNeeRenderer::FromSimple2Json(const std::shared_ptr<ThinlensCamera> &camera,
                                                                  const std::shared_ptr<const SceneData> &scene,
                                                                  const wj::Value &json, RenderContext *ctx) {
    auto r = std::make_shared<PathNeeRenderer>(camera, scene, int(json.number("num_path_vertices", 10)),
                                               int(json.number("num_spp", 16)), ctx);
    r->seed = uint32_t(json.number("seed", 0));
    r->use_mis = json.boolean("use_mis", true);
    // "guiding": {"guide_fraction", "spp_per_pass", "train_steps", "warmup_passes",
    //             "max_guided_bounce", "train_spp", "overlap_training", "learning_rate",
    //             "batch_size", "record_capacity", "avg_decay", "seed",
    //             "inference": "bf16" | "fp8"} (all optional; train_spp > 0 trains on the
    //             first train_spp samples per pixel only and renders the rest in one pass)
    if (const wj::Value *gj = json.find("guiding")) {
        if (gj->type != wj::Value::Object) throw std::runtime_error("\"guiding\" must be an object");
        r->guided = gj->boolean("enabled", true);
        wr_guide_config_default(&r->guide_config);
        r->guide_config.learning_rate = float(gj->number("learning_rate", r->guide_config.learning_rate));
        r->guide_config.batch_size = uint32_t(gj->number("batch_size", r->guide_config.batch_size));
        r->guide_config.record_capacity = uint32_t(gj->number("record_capacity", r->guide_config.record_capacity));
        r->guide_config.seed = uint32_t(gj->number("seed", 0));
        r->guide_config.avg_decay = float(gj->number("avg_decay", r->guide_config.avg_decay));
        r->guide_train_spp = int(gj->number("train_spp", 0));
        r->guide_params.overlap_training = gj->boolean("overlap_training", false) ? 1u : 0u;
        r->guide_params.guide_fraction = float(gj->number("guide_fraction", 0.5));
        r->guide_params.spp_per_pass = uint32_t(gj->number("spp_per_pass", 4));
        r->guide_params.train_steps = uint32_t(gj->number("train_steps", 4));
        r->guide_params.warmup_passes = uint32_t(gj->number("warmup_passes", 1));
        r->guide_params.max_guided_bounce = uint32_t(gj->number("max_guided_bounce", 0));
        const std::string inference = gj->string("inference", "bf16");
        if (inference != "bf16" && inference != "fp8") throw std::runtime_error("guiding.inference must be bf16 or fp8");
        r->guide_params.fp8_inference = inference == "fp8" ? 1 : 0;
        if (!(r->guide_params.guide_fraction >= 0.f && r->guide_params.guide_fraction < 1.f))
            throw std::runtime_error("guide_fraction must be in [0, 1)");
    }
    return r;
}

PathNeeRenderer::PathNeeRenderer(std::shared_ptr<ThinlensCamera> c, std::shared_ptr<const SceneData> s, int n, int spp,
                                 RenderContext *x)
    : camera(std::move(c)), scene(std::move(s)), ctx(x), num_path_vertices(n), num_spp(spp) {}

// render (pathnextevent.h:152-191): film scales 1/spp and 1/(spp W H), then
// num_spp samples per pixel.
void PathNeeRenderer::render() {
    Film &film = *camera->film;
    film.base_scale = 1.0 / double(num_spp);
    film.splat_scale = 1.0 / double(num_spp) / (double(film.W) * double(film.H));
    if (num_path_vertices <= 1 || num_spp <= 0) return;
    const wr_scene_desc d = scene->desc(camera->desc);
    wr_scene *sc = nullptr;
    wr_film *f = nullptr;
    wr_guide *g = nullptr;
    check(wr_scene_create(ctx->get(), &d, &sc));
    int rc = camera->has_medium ? wr_scene_set_medium(sc, &camera->medium) : WR_OK;
    if (rc == WR_OK && !scene->env_rgb.empty()) rc = wr_scene_set_envmap(sc, scene->env_rgb.data(), scene->env_w, scene->env_h);
    if (rc == WR_OK) rc = wr_film_create(sc, &f);
    if (rc == WR_OK && guided) rc = wr_guide_create(sc, &guide_config, &g);
    if (rc == WR_OK) {
        wr_render_params p{};
        p.spp = num_spp;
        p.num_path_vertices = num_path_vertices;
        p.seed = seed;
        p.use_mis = use_mis ? 1 : 0;
        if (!guided) {
            rc = wr_render_nee(sc, f, &p, &stats);
        } else if (guide_train_spp <= 0 || guide_train_spp >= num_spp) {
            rc = wr_render_guided(sc, f, g, &p, &guide_params, &stats, &guide_stats);
        } else {
            // train on the first guide_train_spp samples, then render the rest with the trained network
            p.spp = guide_train_spp;
            rc = wr_render_guided(sc, f, g, &p, &guide_params, &stats, &guide_stats);
            if (rc == WR_OK) {
                wr_guided_params rest = guide_params;
                rest.spp_per_pass = uint32_t(num_spp - guide_train_spp);
                rest.train_steps = 0;
                rest.warmup_passes = 0;
                wr_render_stats st2{};
                wr_guide_stats gs2{};
                p.first_sample = guide_train_spp;
                p.spp = num_spp - guide_train_spp;
                rc = wr_render_guided(sc, f, g, &p, &rest, &st2, &gs2);
                stats.paths += st2.paths;
                stats.extend_rays += st2.extend_rays;
                stats.shadow_rays += st2.shadow_rays;
                stats.wall_ms += st2.wall_ms;
            }
        }
    }
    if (rc == WR_OK) rc = wr_film_read(f, film.base.data(), film.splat.data());
    const std::string err = rc == WR_OK ? "" : wr_last_error();
    if (g) wr_guide_destroy(g);
    if (f) wr_film_destroy(f);
    wr_scene_destroy(sc);
    if (rc != WR_OK) throw std::runtime_error(err);
}

// A render of one entry point on a fresh GPU scene and film, read back into
// the host film.
template <typename Fn>
static void render_on_gpu(const ThinlensCamera &camera, const SceneData &scene, RenderContext *ctx, Fn &&fn) {
    const wr_scene_desc d = scene.desc(camera.desc);
    wr_scene *sc = nullptr;
    wr_film *f = nullptr;
    check(wr_scene_create(ctx->get(), &d, &sc));
    int rc = camera.has_medium ? wr_scene_set_medium(sc, &camera.medium) : WR_OK;
    if (rc == WR_OK && !scene.env_rgb.empty()) rc = wr_scene_set_envmap(sc, scene.env_rgb.data(), scene.env_w, scene.env_h);
    if (rc == WR_OK) rc = wr_film_create(sc, &f);
    if (rc == WR_OK) rc = fn(sc, f);
    if (rc == WR_OK) rc = wr_film_read(f, camera.film->base.data(), camera.film->splat.data());
    const std::string err = rc == WR_OK ? "" : wr_last_error();
    if (f) wr_film_destroy(f);
    wr_scene_destroy(sc);
    if (rc != WR_OK) throw std::runtime_error(err);
}

std::shared_ptr<PrimitivePathRenderer> PrimitivePathRenderer::FromSimple2Json(
    const std::shared_ptr<ThinlensCamera> &camera, const std::shared_ptr<const SceneData> &scene,
    const wj::Value &json, RenderContext *ctx) {
    auto r = std::make_shared<PrimitivePathRenderer>();
    r->camera = camera;
    r->scene = scene;
    r->ctx = ctx;
    r->num_path_vertices = int(json.number("num_path_vertices", 10));
    r->num_spp = int(json.number("num_spp", 16));
    r->seed = uint32_t(json.number("seed", 0));
    return r;
}

// render (primitivepath.h:60-94): scales 1/spp and 1/(spp W H)
void PrimitivePathRenderer::render() {
    Film &film = *camera->film;
    film.base_scale = 1.0 / double(num_spp);
    film.splat_scale = 1.0 / double(num_spp) / (double(film.W) * double(film.H));
    if (num_path_vertices <= 1 || num_spp <= 0) return;
    render_on_gpu(*camera, *scene, ctx, [&](wr_scene *sc, wr_film *f) {
        wr_render_params p{};
        p.spp = num_spp;
        p.num_path_vertices = num_path_vertices;
        p.seed = seed;
        return wr_render_primitive(sc, f, &p, &stats);
    });
}

std::shared_ptr<BidirPathRenderer> BidirPathRenderer::FromSimple2Json(
    const std::shared_ptr<ThinlensCamera> &camera, const std::shared_ptr<const SceneData> &scene,
    const wj::Value &json, RenderContext *ctx) {
    auto r = std::make_shared<BidirPathRenderer>();
    r->camera = camera;
    r->scene = scene;
    r->ctx = ctx;
    r->num_path_vertices = int(json.number("num_path_vertices", 10));
    r->num_spp = int(json.number("num_spp", 16));
    r->seed = uint32_t(json.number("seed", 0));
    return r;
}

// render (bidirpath.h:316-395): base scaled 1/spp, atomic buffer 1/(spp W H)
void BidirPathRenderer::render() {
    Film &film = *camera->film;
    film.base_scale = 1.0 / double(num_spp);
    film.splat_scale = 1.0 / double(num_spp) / (double(film.W) * double(film.H));
    if (num_path_vertices <= 1 || num_spp <= 0) return;
    render_on_gpu(*camera, *scene, ctx, [&](wr_scene *sc, wr_film *f) {
        wr_render_params p{};
        p.spp = num_spp;
        p.num_path_vertices = num_path_vertices;
        p.seed = seed;
        p.use_mis = 1;
        return wr_render_bdpt(sc, f, &p, &stats);
    });
}

std::shared_ptr<LightSplattingRenderer> LightSplattingRenderer::FromSimple2Json(
    const std::shared_ptr<ThinlensCamera> &camera, const std::shared_ptr<const SceneData> &scene,
    const wj::Value &json, RenderContext *ctx) {
    auto r = std::make_shared<LightSplattingRenderer>();
    r->camera = camera;
    r->scene = scene;
    r->ctx = ctx;
    r->num_path_vertices = int(json.number("num_path_vertices", 10));
    r->num_spp = int(json.number("num_spp", 16));
    r->seed = uint32_t(json.number("seed", 0));
    return r;
}

// render (lightsplatting.h:37-113): atomic buffer only, scale 1/(spp W H)
void LightSplattingRenderer::render() {
    Film &film = *camera->film;
    film.base_scale = 0.0;
    film.splat_scale = 1.0 / double(num_spp) / (double(film.W) * double(film.H));
    if (num_path_vertices < 2 || num_spp <= 0) return;
    render_on_gpu(*camera, *scene, ctx, [&](wr_scene *sc, wr_film *f) {
        wr_render_params p{};
        p.spp = num_spp;
        p.num_path_vertices = num_path_vertices;
        p.seed = seed;
        return wr_render_light(sc, f, &p, &stats);
    });
}

std::shared_ptr<VisualizeGeometryNormal> VisualizeGeometryNormal::FromSimple2Json(
    const std::shared_ptr<ThinlensCamera> &camera, const std::shared_ptr<const SceneData> &scene,
    const wj::Value &json, RenderContext *ctx) {
    auto r = std::make_shared<VisualizeGeometryNormal>();
    r->camera = camera;
    r->scene = scene;
    r->ctx = ctx;
    r->num_spp = int(json.number("num_spp", 16));
    r->seed = uint32_t(json.number("seed", 0));
    return r;
}

// render (visualizegeomnormal.h:32-74): the reference adds the per-pixel mean
// at scale 1; the device sums the samples, so the host scales by 1/spp.
void VisualizeGeometryNormal::render() {
    Film &film = *camera->film;
    film.base_scale = num_spp > 0 ? 1.0 / double(num_spp) : 1.0;
    film.splat_scale = 1.0;
    if (num_spp <= 0) return;
    render_on_gpu(*camera, *scene, ctx, [&](wr_scene *sc, wr_film *f) {
        wr_render_params p{};
        p.spp = num_spp;
        p.seed = seed;
        return wr_render_visualize(sc, f, &p, &stats);
    });
}

void LoadPfm(const std::string &path, std::vector<float> *rgb, uint32_t *width, uint32_t *height) {
    std::ifstream is(path, std::ios::binary);
    if (!is.is_open()) throw std::runtime_error("cannot open envmap " + path);
    std::string tag, scale;
    long w = -1, h = -1;
    is >> tag >> w >> h >> scale;
    int c = is.get();
    if (c == '\r') c = is.get();
    if ((tag != "PF" && tag != "Pf") || w <= 0 || h <= 0 || !is) throw std::runtime_error("not a PFM image: " + path);
    const int ch = tag == "PF" ? 3 : 1;
    std::vector<float> row(size_t(ch) * size_t(w));
    rgb->assign(3 * size_t(w) * size_t(h), 0.f);
    for (long y = h; y > 0; y--) {  // rows bottom-up: the first row read is y = h - 1
        is.read(reinterpret_cast<char *>(row.data()), std::streamsize(sizeof(float) * row.size()));
        if (!is) throw std::runtime_error("truncated PFM image: " + path);
        for (long x = 0; x < w; x++)
            for (int k = 0; k < 3; k++) (*rgb)[3 * (size_t(y - 1) * size_t(w) + size_t(x)) + k] = row[size_t(x) * ch + (ch == 3 ? k : 0)];
    }
    *width = uint32_t(w);
    *height = uint32_t(h);
}

// ------------------------------------------------------------------ scene JSON
namespace {

template <typename T>
void dump_array(const std::string &dir, const char *name, const std::vector<T> &v) {
    std::ofstream os(dir + "/" + name, std::ios::binary);
    if (!os.is_open()) throw std::runtime_error(std::string("cannot write ") + dir + "/" + name);
    os.write(reinterpret_cast<const char *>(v.data()), sizeof(T) * v.size());
}

// integrator names the reference's dispatch knows (simple2parser.h:176-207)
const char *kOtherIntegrators[] = {"photon_map", "diffuse_prt", "glossy_prt2002"};

}  // namespace

bool ParseMedium(const wj::Value &json, wr_medium *out) {
    const std::string type = json.string("type", "");
    if (type == "none") return false;
    if (type != "homogeneous") throw std::runtime_error("unknown medium");
    const auto a = json.vec3("sigma_a"), s = json.vec3("sigma_s");
    for (int c = 0; c < 3; c++) {
        out->sigma_a[c] = float(a[c]);
        out->sigma_s[c] = float(s[c]);
    }
    out->g = float(json.number("g", 0.0));
    if (!(out->g > -1.f && out->g < 1.f)) throw std::runtime_error("medium \"g\" must be in (-1, 1)");
    for (int c = 0; c < 3; c++)
        if (!(out->sigma_a[c] >= 0.f && out->sigma_s[c] >= 0.f))
            throw std::runtime_error("medium coefficients must be >= 0");
    return true;
}

void Simple2Parser::Parse(const std::string &json_path, const ParseOptions &opt) {
    std::ifstream is(json_path);
    if (!is.is_open()) throw std::runtime_error("file not found" + json_path);
    std::stringstream ss;
    ss << is.rdbuf();
    Parse(wj::parse(ss.str()), dir_of(json_path), opt);
}

void Simple2Parser::Parse(const wj::Value &j, const std::string &root_dir, const ParseOptions &opt) {
    if (j.type != wj::Value::Object) throw std::runtime_error("scene json must be an object");

    if (j.has("render2d")) throw std::runtime_error("2-D rendering is not part of this build");
    auto scene = std::make_shared<SceneData>();
    if (const wj::Value *meshes = j.find("meshes")) {
        if (meshes->type != wj::Value::Array) throw std::runtime_error("\"meshes\" must be an array");
        for (const wj::Value &m : meshes->arr) {
            if (m.boolean("phong_tessellation", false)) throw std::runtime_error("phong_tessellation is not part of this build");
            LoadObj(root_dir + "/" + m.string("path", ""), m.boolean("mitsuba_srgb", false), scene.get());
        }
    }
    const std::string envmap = j.string("envmap", "");
    if (!envmap.empty()) {
        const size_t dot = envmap.find_last_of('.');
        const std::string ext = dot == std::string::npos ? "" : envmap.substr(dot);
        if (ext != ".pfm") throw std::runtime_error("envmap \"" + envmap + "\": only .pfm images are part of this build");
        LoadPfm(root_dir + "/" + envmap, &scene->env_rgb, &scene->env_w, &scene->env_h);
        if (scene->env_w != 2 * scene->env_h) throw std::runtime_error("envmap has wrong size");
    }
    if (scene->lights.empty() && scene->env_rgb.empty()) throw std::runtime_error("scene has no lights");
    if (!j.has("render")) return;
    auto camera = ThinlensCamera::FromSimple2ParseJson(j.at("camera"));
    if (const wj::Value *med = j.find("medium")) camera->has_medium = ParseMedium(*med, &camera->medium);
    const wr_scene_desc d = scene->desc(camera->desc);
    if (wr_scene_validate(&d) != WR_OK) throw std::runtime_error(wr_last_error());
    if (!opt.dump_dir.empty()) {
        dump_array(opt.dump_dir, "verts.f32", scene->verts);
        dump_array(opt.dump_dir, "tris.u32", scene->tris);
        dump_array(opt.dump_dir, "tri_geom.u32", scene->tri_geom);
        dump_array(opt.dump_dir, "geom_material.i32", scene->geom_material);
        dump_array(opt.dump_dir, "geom_light.i32", scene->geom_light);
        dump_array(opt.dump_dir, "materials.bin", scene->materials);
        dump_array(opt.dump_dir, "lights.bin", scene->lights);
        std::vector<wr_camera> cam{camera->desc};
        dump_array(opt.dump_dir, "camera.bin", cam);
    }
    RenderContext ctx;
    const wj::Value &renders = j.at("render");
    if (renders.type != wj::Value::Array) throw std::runtime_error("\"render\" must be an array");
    for (const wj::Value &rj : renders.arr) {
        const std::string name = rj.string("integrator", "");
        std::shared_ptr<Renderer> renderer;
        if (name == "nextevent_path_tracer") {
            renderer = PathNeeRenderer::FromSimple2Json(camera, scene, rj, &ctx);
        } else if (name == "primitive_path_tracer") {
            renderer = PrimitivePathRenderer::FromSimple2Json(camera, scene, rj, &ctx);
        } else if (name == "light_splatting") {
            renderer = LightSplattingRenderer::FromSimple2Json(camera, scene, rj, &ctx);
        } else if (name == "visualize") {
            renderer = VisualizeGeometryNormal::FromSimple2Json(camera, scene, rj, &ctx);
        } else if (name == "bidirectional_path_tracer") {
            renderer = BidirPathRenderer::FromSimple2Json(camera, scene, rj, &ctx);
        } else {
            for (const char *other : kOtherIntegrators)
                if (name == other) throw std::runtime_error("integrator \"" + name + "\" is not part of this build");
            throw std::runtime_error("unknown integrator \"" + name + "\"");
        }
        if (opt.check_only) continue;
        // each render entry starts from an empty film: the reference's renderers
        // request fresh, zeroed buffers (film.h:136-150, pathnextevent.h:154-155)
        camera->film = std::make_shared<Film>(camera->desc.width, camera->desc.height);
        renderer->render();
        const std::string output = rj.string("output", "");
        if (!output.empty()) camera->film->save_pfm(root_dir + "/" + output);
        if (auto *nee = dynamic_cast<PathNeeRenderer *>(renderer.get())) {
            std::printf("{\"integrator\": \"%s\", \"output\": \"%s\", \"spp\": %d, \"paths\": %llu, \"extend_rays\": %llu, "
                        "\"shadow_rays\": %llu, \"ms\": %.3f, \"guided\": %s, \"train_steps\": %llu}\n",
                        name.c_str(), output.c_str(), nee->num_spp, (unsigned long long)nee->stats.paths,
                        (unsigned long long)nee->stats.extend_rays, (unsigned long long)nee->stats.shadow_rays,
                        nee->stats.wall_ms, nee->guided ? "true" : "false",
                        (unsigned long long)nee->guide_stats.train_steps);
        } else {
            const wr_render_stats *st = nullptr;
            int spp = 0;
            if (auto *pr = dynamic_cast<PrimitivePathRenderer *>(renderer.get())) { st = &pr->stats; spp = pr->num_spp; }
            if (auto *vz = dynamic_cast<VisualizeGeometryNormal *>(renderer.get())) { st = &vz->stats; spp = vz->num_spp; }
            if (auto *ls = dynamic_cast<LightSplattingRenderer *>(renderer.get())) { st = &ls->stats; spp = ls->num_spp; }
            if (auto *bd = dynamic_cast<BidirPathRenderer *>(renderer.get())) { st = &bd->stats; spp = bd->num_spp; }
            if (st)
                std::printf("{\"integrator\": \"%s\", \"output\": \"%s\", \"spp\": %d, \"paths\": %llu, "
                            "\"extend_rays\": %llu, \"ms\": %.3f}\n",
                            name.c_str(), output.c_str(), spp, (unsigned long long)st->paths,
                            (unsigned long long)st->extend_rays, st->wall_ms);
        }
    }
}

}  // namespace wurst
