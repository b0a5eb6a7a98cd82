// ref_vectors.cpp -- golden-vector generator built from the reference's own
// headers (the per-call classes of the `nextevent_path_tracer` path).
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile into oracle/_ref/ and
// run by tests/golden/make_golden.py in the container that holds the reference
// checkout; its JSON output is committed as tests/golden/ref_vectors.json.
// It includes, and so exercises, these reference classes unchanged:
//   CoordFrame3 (src/common/math/coordframe.h:41-91), Mapping (src/common/mapping.h),
//   CdfTable (src/common/cdftable.h), LambertDiffuseBsdf, OrenNayarBsdf, PhongBsdf,
//   MixedBsdf, TwoSidedBsdf (src/bsdf/*.h), DiffuseAreaLight
//   (src/light/diffusearealight.h) on TriangleMeshGeometry
//   (src/geometry/trianglemesh.h), ThinlensCamera (src/camera/thinlens.h),
//   EquiRectRep (src/sphericalrepresentation/equirectangular.h: the envmap's
//   bilinear lookup, pixel CDF, pdf and sampling; Fimage, src/common/floatimage.h).
// The light-selection composition mirrors GeneralSubpathSampler::createLightSubpath
// (src/pathsampler/generalsubpathsampler.h:203-250), which cannot be
// instantiated without Embree.  Scene inputs (light meshes, cameras) are exactly
// representable in FP32 because the oracle and the GPU take FP32 scene data.
#include "common/util/json.h"
#include "camera/thinlens.h"
#include "bsdf/lambertdiffuse.h"
#include "bsdf/orennayar.h"
#include "bsdf/phong.h"
#include "bsdf/mixed.h"
#include "bsdf/twosided.h"
#include "light/diffusearealight.h"
#include "geometry/trianglemesh.h"
#include "common/cdftable.h"
#include "common/mapping.h"
#include "sphericalrepresentation/equirectangular.h"

#include <cstdio>
#include <cstdint>

static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static double rnd() {  // xorshift64*, inputs only
    g_state ^= g_state >> 12; g_state ^= g_state << 25; g_state ^= g_state >> 27;
    return double((g_state * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0);
}
static Vec3 rnd_dir() {
    for (;;) {
        Vec3 v(2 * rnd() - 1, 2 * rnd() - 1, 2 * rnd() - 1);
        double l2 = Math::Length2(v);
        if (l2 > 1e-4 && l2 <= 1.0) return v / std::sqrt(l2);
    }
}
static void pv(const char *k, const Vec3 &v, bool comma = true) {
    printf("\"%s\":[%.17g,%.17g,%.17g]%s", k, v[0], v[1], v[2], comma ? "," : "");
}
static void pv2(const char *k, const Vec2 &v, bool comma = true) {
    printf("\"%s\":[%.17g,%.17g]%s", k, v[0], v[1], comma ? "," : "");
}
static void pd(const char *k, double v, bool comma = true) { printf("\"%s\":%.17g%s", k, v, comma ? "," : ""); }

struct Mat {
    int type, twosided;
    double rd[3], rs[3], sigma, n;
    shared_ptr<const Bsdf> bsdf;
};

static Mat make_mat(int type, int twosided, Vec3 rd, Vec3 rs, double sigma, double n) {
    Mat m;
    m.type = type; m.twosided = twosided;
    for (int i = 0; i < 3; i++) { m.rd[i] = rd[i]; m.rs[i] = rs[i]; }
    m.sigma = sigma; m.n = n;
    auto tex = [](Vec3 c) { return make_shared<const ConstantTexture<Spectrum>>(Spectrum(c[0], c[1], c[2])); };
    shared_ptr<const Bsdf> lam = make_shared<LambertDiffuseBsdf>(tex(rd));
    shared_ptr<const Bsdf> on = make_shared<OrenNayarBsdf>(tex(rd), tex(Vec3(sigma)));
    shared_ptr<const Bsdf> ph = make_shared<PhongBsdf>(tex(rs), make_shared<const ConstantTexture<double>>(n));
    shared_ptr<const Bsdf> inner;
    if (type == 0) inner = lam;
    else if (type == 1) inner = on;
    else if (type == 2) inner = ph;
    else { std::vector<shared_ptr<const Bsdf>> v{lam, ph}; inner = make_shared<MixedBsdf>(v); }
    m.bsdf = twosided ? shared_ptr<const Bsdf>(make_shared<TwoSidedBsdf>(inner)) : inner;
    return m;
}

static void print_mat(const Mat &m) {
    printf("\"mat\":{\"type\":%d,\"twosided\":%d,\"rho_d\":[%.17g,%.17g,%.17g],\"rho_s\":[%.17g,%.17g,%.17g],\"sigma\":%.17g,\"exponent\":%.17g},",
           m.type, m.twosided, m.rd[0], m.rd[1], m.rd[2], m.rs[0], m.rs[1], m.rs[2], m.sigma, m.n);
}

int main() {
    printf("{\n");
    // ---- ONB
    printf("\"onb\":[\n");
    for (int i = 0; i < 48; i++) {
        Vec3 n = rnd_dir();
        if (i == 0) n = Vec3(0, 1, 0);
        if (i == 1) n = Vec3(0, -1, 0);
        if (i == 2) n = Math::Normalize(Vec3(1e-9, -1, 1e-9));
        CoordFrame3 f(n);
        printf("{"); pv("n", n); pv("x", f.x); pv("z", f.z, false); printf("}%s\n", i == 47 ? "" : ",");
    }
    printf("],\n");
    // ---- mappings
    printf("\"mapping\":[\n");
    for (int i = 0; i < 48; i++) {
        Vec2 u(rnd(), rnd());
        if (i == 0) u = Vec2(0.5, 0.5);
        if (i == 1) u = Vec2(0.0, 0.0);
        printf("{"); pv2("u", u);
        pv("cosine_hemisphere", Mapping::CosineWeightedHemisphereFromSquare(u));
        pv2("barycentric", Mapping::BarycentricFromSquare(u));
        pv2("disk", Mapping::DiskFromSquare(u), false);
        printf("}%s\n", i == 47 ? "" : ",");
    }
    printf("],\n");
    // ---- CDF tables
    printf("\"cdf\":[\n");
    {
        std::vector<std::vector<double>> tables = {{3, 1, 0, 2, 5}, {1}, {0.25, 0.75}, {2, 2, 2}};
        bool first = true;
        for (auto &w : tables) {
            CdfTable t(w);
            for (int i = 0; i < 12; i++) {
                double u = rnd();
                if (i == 0) u = 0.0;
                if (i == 1) u = t.mCdfs[0];
                double pdf = -1, rem = -1;
                int idx = t.sample(&pdf, &rem, u);
                printf("%s{\"w\":[", first ? "" : ",\n");
                first = false;
                for (size_t k = 0; k < w.size(); k++) printf("%.17g%s", w[k], k + 1 < w.size() ? "," : "");
                printf("],"); pd("u", u); printf("\"index\":%d,", idx); pd("pdf", pdf); pd("remapped", rem, false); printf("}");
            }
        }
        printf("\n],\n");
    }
    // ---- BSDFs
    std::vector<Mat> mats;
    for (int ts = 0; ts < 2; ts++) {
        mats.push_back(make_mat(0, ts, Vec3(0.63, 0.065, 0.05), Vec3(0), 0, 0));
        mats.push_back(make_mat(1, ts, Vec3(0.5, 0.7, 0.2), Vec3(0), 0.35, 0));
        mats.push_back(make_mat(2, ts, Vec3(0), Vec3(0.8, 0.8, 0.8), 0, 37.5));
        mats.push_back(make_mat(2, ts, Vec3(0), Vec3(0.3, 0.5, 0.9), 0, 900.0));
        mats.push_back(make_mat(3, ts, Vec3(0.4, 0.3, 0.2), Vec3(0.5, 0.5, 0.5), 0, 12.0));
    }
    SurfaceVertex sv;
    sv.mTextureCoord = Vec2(0.25, 0.75);
    printf("\"bsdf_eval\":[\n");
    {
        bool first = true;
        for (auto &m : mats)
            for (int i = 0; i < 40; i++) {
                Vec3 wi = rnd_dir(), wo = rnd_dir();
                if (m.type == 2 && i % 2 == 0) {  // near the mirror lobe so Phong is non-zero
                    Vec3 r(-wi[0], wi[1], -wi[2]);
                    wo = Math::Normalize(r + 0.05 * rnd_dir());
                }
                Spectrum f = m.bsdf->evalBsdf(sv, wi, wo);
                double pdf = m.bsdf->evalPdfW(sv, wi, wo);
                printf("%s{", first ? "" : ",\n");
                first = false;
                print_mat(m); pv("wi", wi); pv("wo", wo);
                printf("\"f\":[%.17g,%.17g,%.17g],", f[0], f[1], f[2]);
                pd("pdf", pdf, false); printf("}");
            }
        printf("\n],\n");
    }
    printf("\"bsdf_sample\":[\n");
    {
        bool first = true;
        for (auto &m : mats)
            for (int i = 0; i < 40; i++) {
                Vec3 wi = rnd_dir();
                if (!m.twosided && wi[1] <= 0) wi = Vec3(wi[0], -wi[1], wi[2]);
                Vec2 u(rnd(), rnd());
                Vec3 wo(0.0);
                double pdf = 0.0;
                Spectrum w = m.bsdf->sampleBsdfCosByPdf(&wo, &pdf, sv, wi, u);
                printf("%s{", first ? "" : ",\n");
                first = false;
                print_mat(m); pv("wi", wi); pv2("u", u); pv("wo", wo);
                printf("\"w\":[%.17g,%.17g,%.17g],", w[0], w[1], w[2]);
                pd("pdf", pdf, false); printf("}");
            }
        printf("\n],\n");
    }
    // ---- area lights: two meshes, selection by CdfTable over cdfWeight()
    {
        auto mesh_a = make_shared<TriangleMeshGeometry>();
        mesh_a->mPositions = {Vec3(-0.25, 0.9990234375, -0.25), Vec3(0.25, 0.9990234375, -0.25), Vec3(0.25, 0.9990234375, 0.25), Vec3(-0.25, 0.9990234375, 0.25)};
        mesh_a->mTriangles = {Ivec3(0, 1, 2), Ivec3(0, 2, 3)};
        mesh_a->mTextureCoords = {Vec2(0, 0), Vec2(1, 0), Vec2(1, 1), Vec2(0, 1)};
        auto mesh_b = make_shared<TriangleMeshGeometry>();
        mesh_b->mPositions = {Vec3(2, 8, 0), Vec3(3.5, 8, 0.25), Vec3(2.375, 8.25, 1.125), Vec3(4, 7.5, 1.875)};
        mesh_b->mTriangles = {Ivec3(0, 1, 2), Ivec3(1, 3, 2), Ivec3(0, 2, 3)};
        mesh_b->mTextureCoords = {Vec2(0, 0), Vec2(1, 0), Vec2(1, 1), Vec2(0, 1)};
        auto la = make_shared<DiffuseAreaLight>(Spectrum(17.0, 17.0, 17.0));
        auto lb = make_shared<DiffuseAreaLight>(Spectrum(20.0, 5.0, 1.0));
        mesh_a->mAreaLight = la; mesh_b->mAreaLight = lb;
        mesh_a->precomputeForSampling(); mesh_b->precomputeForSampling();
        std::vector<shared_ptr<TriangleMeshGeometry>> meshes{mesh_a, mesh_b};
        std::vector<double> weights{la->cdfWeight(), lb->cdfWeight()};
        CdfTable lightCdf(weights);
        printf("\"light_meshes\":[");
        for (size_t m = 0; m < meshes.size(); m++) {
            printf("%s{\"verts\":[", m ? "," : "");
            for (size_t k = 0; k < meshes[m]->mPositions.size(); k++) {
                auto p = meshes[m]->mPositions[k];
                printf("%s%.17g,%.17g,%.17g", k ? "," : "", p[0], p[1], p[2]);
            }
            printf("],\"tris\":[");
            for (size_t k = 0; k < meshes[m]->mTriangles.size(); k++) {
                auto t = meshes[m]->mTriangles[k];
                printf("%s%d,%d,%d", k ? "," : "", t[0], t[1], t[2]);
            }
            Spectrum L = (m == 0) ? Spectrum(17.0, 17.0, 17.0) : Spectrum(20.0, 5.0, 1.0);
            printf("],\"L\":[%.17g,%.17g,%.17g],\"area\":%.17g,\"cdf_weight\":%.17g}", L[0], L[1], L[2],
                   meshes[m]->area(), weights[m]);
        }
        printf("],\n\"light_sample\":[\n");
        for (int i = 0; i < 48; i++) {
            Vec2 u(rnd(), rnd());
            Vec2 sample = u;
            double pdfLight;
            int idx = lightCdf.sample(&pdfLight, &sample[0], sample[0]);
            LightVertex lv;
            lv.mGeometry = meshes[idx].get();
            double pdfPos;
            Spectrum c0 = meshes[idx]->mAreaLight->sampleLe0ByPdf(&lv, &pdfPos, sample) / pdfLight;
            printf("{"); pv2("u", u); printf("\"light\":%d,", idx); pv("pos", lv.mPosition); pv("normal", lv.mGeometryNormal);
            printf("\"contrib\":[%.17g,%.17g,%.17g],", c0[0], c0[1], c0[2]);
            pd("pdfA", pdfLight * pdfPos);
            Vec3 w = rnd_dir();
            Spectrum le1 = lv.mLight->evalLe1(lv, lv.mCoordFrame.toLocal(w));
            pv("w", w); pd("le1", le1[0]); pd("pdf_le0", lightCdf.mPdfs[idx] * lv.mLight->evalPdfLe0(lv), false);
            printf("}%s\n", i == 47 ? "" : ",");
        }
        printf("],\n");
    }
    // ---- thin-lens camera (pinhole and lens)
    printf("\"camera\":[\n");
    {
        struct Cam { Vec3 pos, at, up; double fovy; int w, h; double ap, focal; };
        std::vector<Cam> cams = {{Vec3(0, 0, 3.75), Vec3(0, 0, 0), Vec3(0, 1, 0), 40.0, 256, 256, 0.0, 1.0},
                                 {Vec3(0, 2, 18), Vec3(0, 0, 0), Vec3(0, 1, 0), 40.0, 1920, 1080, 0.0, 1.0},
                                 {Vec3(4, 1, -2), Vec3(-1, 1, 0), Vec3(0, 1, 0), 50.0, 64, 32, 0.125, 3.0}};
        bool first = true;
        for (auto &c : cams) {
            ThinlensCamera cam(c.pos, c.at, c.up, c.fovy * Math::Pi / 180.0, Ivec2(c.w, c.h), c.ap, c.focal);
            for (int i = 0; i < 16; i++) {
                Ivec2 pixel(int(rnd() * c.w), int(rnd() * c.h));
                Vec2 ul(rnd(), rnd()), up(rnd(), rnd());
                CameraVertex cv;
                double pdfA = 0;
                cam.sampleWe0ByPdf(&cv, &pdfA, ul);
                Vec3 dir; Vec2 raster; double pdfW = 0;
                cam.samplePerPixelWe1CosByPdf(&dir, &raster, &pdfW, cv, pixel, up);
                Vec3 w = Math::Normalize(dir + 0.2 * rnd_dir());
                Vec2 r2(-1, -1);
                Spectrum we1 = cam.evalWe1(&r2, cv, w);
                printf("%s{\"cam\":{\"pos\":[%.17g,%.17g,%.17g],\"lookat\":[%.17g,%.17g,%.17g],\"up\":[%.17g,%.17g,%.17g],"
                       "\"fovy\":%.17g,\"width\":%d,\"height\":%d,\"aperture\":%.17g,\"focal\":%.17g},",
                       first ? "" : ",\n", c.pos[0], c.pos[1], c.pos[2], c.at[0], c.at[1], c.at[2], c.up[0], c.up[1], c.up[2],
                       c.fovy, c.w, c.h, c.ap, c.focal);
                first = false;
                printf("\"pixel\":[%d,%d],", pixel[0], pixel[1]);
                pv2("ulens", ul); pv2("upix", up); pv("org", cv.mPosition); pd("pdfA", pdfA); pv("dir", dir); pv2("raster", raster);
                pd("pdfW", pdfW); pd("pdfW_eval", cam.evalPerPixelWe1PdfW(cv, dir)); pv("w", w); pd("we1", we1[0]); pv2("we1_raster", r2, false);
                printf("}");
            }
        }
        printf("\n],\n");
    }
    // ---- equirectangular envmap (EquiRectRep with its pixel CDF, as EnvmapLight uses it)
    printf("\"envmap\":{\n");
    {
        const int W = 8, H = 4;
        std::vector<Spectrum> data(W * H);
        for (auto &t : data) {  // FP32 values: the oracle and the GPU take FP32 images
            const double r = float(4 * rnd()), g = float(4 * rnd()), b = float(4 * rnd());
            t = Spectrum(r, g, b);
        }
        data[5] = Spectrum(0.0);  // a black texel is never sampled
        Fimage<Spectrum> img(W, H, data);
        EquiRectRep<Spectrum> rep(img, true);
        printf("\"width\":%d,\"height\":%d,\"texels\":[", W, H);
        for (int i = 0; i < W * H; i++) printf("%s%.17g,%.17g,%.17g", i ? "," : "", data[i][0], data[i][1], data[i][2]);
        printf("],\n\"lookups\":[\n");
        for (int i = 0; i < 64; i++) {
            Vec3 d = rnd_dir();
            if (i == 0) d = Vec3(0, 1, 0);   // pole: pdf 0
            if (i == 1) d = Vec3(0, -1, 0);  // other pole
            if (i == 2) d = Vec3(-1, 0, 0);  // u = 0.5
            if (i == 3) d = Math::Normalize(Vec3(1, 0.3, -1e-6));  // just before the seam (u -> 1)
            if (i == 4) d = Math::Normalize(Vec3(1, -0.2, 1e-6));  // just after it (u -> 0)
            const Vec2 uv = Mapping::PanoramaFromWorld(d);
            const Spectrum e = rep.eval(d);
            printf("{"); pv("dir", d); pv2("uv", uv); pv("eval", Vec3(e[0], e[1], e[2])); pd("pdf", rep.pdfA(d), false);
            printf("}%s\n", i == 63 ? "" : ",");
        }
        printf("],\n\"samples\":[\n");
        for (int i = 0; i < 64; i++) {
            const Vec2 u(rnd(), rnd());
            Vec3 dir;
            double pdf = 0;
            const Spectrum v = rep.sample(&dir, &pdf, u);
            printf("{"); pv2("u", u); pv("dir", dir); pd("pdf", pdf); pv("value", Vec3(v[0], v[1], v[2]), false);
            printf("}%s\n", i == 63 ? "" : ",");
        }
        printf("],\n\"panorama\":[\n");
        for (int i = 0; i < 32; i++) {
            const Vec2 uv(rnd(), rnd());
            printf("{"); pv2("uv", uv); pv("dir", Mapping::WorldFromPanorama(uv), false); printf("}%s\n", i == 31 ? "" : ",");
        }
        printf("]}\n");
    }
    printf("}\n");
    return 0;
}
