// wurst_host.hpp -- C++ host of the MI355X `nextevent_path_tracer` (and of
// `primitive_path_tracer`, `light_splatting` and `visualize` on the same
// kernels).
//
// Mirrors the reference's scene-JSON entry point (Simple2Parser,
// src/sceneio/simple2parser.h:84-223), its ThinlensCamera JSON reader
// (src/camera/thinlens.h:13-26), the OBJ/MTL material mapping of its Assimp
// loader (src/sceneio/meshio/assimpio.h:61-153) and PathNeeRenderer's
// FromSimple2Json / render contract (src/renderer/pathnextevent.h:17-39,152-191).
// Rendering goes through the C-ABI of include/wurst_amd.h.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "../../include/wurst_amd.h"
#include "wjson.hpp"

namespace wurst {

// Film (src/common/film.h:35-90): FP64 base and splat sums and their scales.
struct Film {
    int W = 0, H = 0;
    std::vector<double> base, splat;  // [H][W][3], row 0 = top of the image
    double base_scale = 1.0, splat_scale = 1.0;
    Film(int w, int h) : W(w), H(h), base(3 * size_t(w) * h, 0.0), splat(3 * size_t(w) * h, 0.0) {}
    // Film::getImage: base * base_scale + splat * splat_scale, in FP32 RGB
    std::vector<float> image() const;
    // FimageIo::SavePfmColor (src/common/floatimage.h:435-462): "PF", FP32, rows bottom-up
    void save_pfm(const std::string &path) const;
};

struct ThinlensCamera {
    wr_camera desc{};
    std::shared_ptr<Film> film;
    // the scene's medium hangs off the camera, as in the reference (simple2parser.h:163-167)
    bool has_medium = false;
    wr_medium medium{};
    static std::shared_ptr<ThinlensCamera> FromSimple2ParseJson(const wj::Value &json);
};

// Triangle meshes in attach order plus their materials and emitters, as the
// C-ABI takes them.
struct SceneData {
    std::vector<float> verts;
    std::vector<uint32_t> tris, tri_geom;
    std::vector<int32_t> geom_material, geom_light;
    std::vector<wr_material> materials;
    std::vector<wr_light> lights;
    // envmap (scene JSON "envmap", simple2parser.h:110-119): RGB [env_h][env_w], row 0 = top
    std::vector<float> env_rgb;
    uint32_t env_w = 0, env_h = 0;
    wr_scene_desc desc(const wr_camera &camera) const;
};

// FimageIo::LoadPfm (src/common/floatimage.h:238-281): "PF" (RGB) or "Pf"
// (grey, replicated), little-endian FP32 rows stored bottom-up.  Other image
// formats of the reference's loader (.hdr, .exr, LDR) are not part of this build.
void LoadPfm(const std::string &path, std::vector<float> *rgb, uint32_t *width, uint32_t *height);

// OBJ + MTL reader with the reference's material mapping:
//   Ke != 0 and Ns == 0 -> DiffuseAreaLight(Ke) (Ke with Ns != 0 is the reference's
//   SimpleMetalAreaLight, not part of this path: rejected),
//   Kd only -> TwoSided(Lambert(Kd)), Ks only -> TwoSided(Phong(Ks, Ns / 4)),
//   both -> TwoSided(Mixed{Lambert, Phong}).
// One geometry per run of faces with the same object and material, in file order.
void LoadObj(const std::string &path, bool srgb_the_mitsuba_way, SceneData *scene);

struct RenderContext {
    wr_device *device = nullptr;
    int device_ordinal = 0;
    ~RenderContext();
    wr_device *get();
};

struct Renderer {
    virtual ~Renderer() = default;
    virtual void render() = 0;
};

// PathNeeRenderer (src/renderer/pathnextevent.h:15-196) on the GPU.
struct PathNeeRenderer : Renderer {
    static std::shared_ptr<PathNeeRenderer> FromSimple2Json(const std::shared_ptr<ThinlensCamera> &camera,
                                                            const std::shared_ptr<const SceneData> &scene,
                                                            const wj::Value &json, RenderContext *ctx);
    PathNeeRenderer(std::shared_ptr<ThinlensCamera> camera, std::shared_ptr<const SceneData> scene, int num_path_vertices,
                    int num_spp, RenderContext *ctx);
    void render() override;

    std::shared_ptr<ThinlensCamera> camera;
    std::shared_ptr<const SceneData> scene;
    RenderContext *ctx;
    int num_path_vertices, num_spp;
    uint32_t seed = 0;
    bool use_mis = true;
    // learned guiding: the optional "guiding" object of the render entry (no reference counterpart)
    bool guided = false;
    wr_guided_params guide_params{};
    int guide_train_spp = 0;  // > 0: train on the first guide_train_spp samples only
    wr_guide_config guide_config{};
    wr_render_stats stats{};
    wr_guide_stats guide_stats{};
};

// PrimitivePathRenderer (src/renderer/primitivepath.h:15-100) on the GPU:
// BSDF-sampled paths that count the emitters they hit; JSON keys
// "num_path_vertices" (10), "num_spp" (16), "seed".
struct PrimitivePathRenderer : Renderer {
    static std::shared_ptr<PrimitivePathRenderer> FromSimple2Json(const std::shared_ptr<ThinlensCamera> &camera,
                                                                  const std::shared_ptr<const SceneData> &scene,
                                                                  const wj::Value &json, RenderContext *ctx);
    void render() override;

    std::shared_ptr<ThinlensCamera> camera;
    std::shared_ptr<const SceneData> scene;
    RenderContext *ctx = nullptr;
    int num_path_vertices = 10, num_spp = 16;
    uint32_t seed = 0;
    wr_render_stats stats{};
};

// LightSplattingRenderer (src/renderer/lightsplatting.h:12-116) on the GPU:
// light subpaths connected to the camera; JSON keys "num_path_vertices" (10),
// "num_spp" (16), "seed".
struct LightSplattingRenderer : Renderer {
    static std::shared_ptr<LightSplattingRenderer> FromSimple2Json(const std::shared_ptr<ThinlensCamera> &camera,
                                                                   const std::shared_ptr<const SceneData> &scene,
                                                                   const wj::Value &json, RenderContext *ctx);
    void render() override;

    std::shared_ptr<ThinlensCamera> camera;
    std::shared_ptr<const SceneData> scene;
    RenderContext *ctx = nullptr;
    int num_path_vertices = 10, num_spp = 16;
    uint32_t seed = 0;
    wr_render_stats stats{};
};

// BidirPathRenderer (src/renderer/bidirpath.h:44-429) on the GPU: light and
// camera subpaths joined by every strategy, balance-heuristic weights; JSON
// keys "num_path_vertices" (10), "num_spp" (16), "seed".
struct BidirPathRenderer : Renderer {
    static std::shared_ptr<BidirPathRenderer> FromSimple2Json(const std::shared_ptr<ThinlensCamera> &camera,
                                                              const std::shared_ptr<const SceneData> &scene,
                                                              const wj::Value &json, RenderContext *ctx);
    void render() override;

    std::shared_ptr<ThinlensCamera> camera;
    std::shared_ptr<const SceneData> scene;
    RenderContext *ctx = nullptr;
    int num_path_vertices = 10, num_spp = 16;
    uint32_t seed = 0;
    wr_render_stats stats{};
};

// VisualizeGeometryNormal (src/renderer/visualizegeomnormal.h:14-76) on the
// GPU: the primary hit's geometric normal as RGB, 1 on emitters; JSON keys
// "num_spp" (16), "seed".
struct VisualizeGeometryNormal : Renderer {
    static std::shared_ptr<VisualizeGeometryNormal> FromSimple2Json(const std::shared_ptr<ThinlensCamera> &camera,
                                                                    const std::shared_ptr<const SceneData> &scene,
                                                                    const wj::Value &json, RenderContext *ctx);
    void render() override;

    std::shared_ptr<ThinlensCamera> camera;
    std::shared_ptr<const SceneData> scene;
    RenderContext *ctx = nullptr;
    int num_spp = 16;
    uint32_t seed = 0;
    wr_render_stats stats{};
};

struct ParseOptions {
    bool check_only = false;     // parse and validate; no GPU
    std::string dump_dir;        // write the flattened scene arrays here
};

// Simple2Parser::ParseMedium (simple2parser.h:70-82): "homogeneous" with
// RgbFromJson "sigma_a", "sigma_s" (homogeneous.h:11-20), plus an optional HG
// asymmetry "g" (default 0: the reference's isotropic phase); "none"; anything
// else throws "unknown medium".
bool ParseMedium(const wj::Value &json, wr_medium *out);

// Simple2Parser::Parse (src/sceneio/simple2parser.h:84-223).  Throws
// std::runtime_error on malformed input or unsupported features.
struct Simple2Parser {
    static void Parse(const std::string &json_path, const ParseOptions &opt = {});
    static void Parse(const wj::Value &json, const std::string &root_dir, const ParseOptions &opt = {});
};

}  // namespace wurst
