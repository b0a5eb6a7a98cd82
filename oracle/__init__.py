"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_VECTORS_BIN = os.path.join(_HERE, "_ref", "ref_vectors")


class OrcMaterial(C.Structure):
    _fields_ = [("type", C.c_int32), ("twosided", C.c_int32), ("rho_d", C.c_float * 3),
                ("rho_s", C.c_float * 3), ("sigma", C.c_float), ("exponent", C.c_float)]


class OrcCamera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("lookat", C.c_float * 3), ("up", C.c_float * 3),
                ("fovy_deg", C.c_float), ("width", C.c_int32), ("height", C.c_int32),
                ("aperture_radius", C.c_float), ("focal_distance", C.c_float)]


class OrcSceneDesc(C.Structure):
    _fields_ = [("verts", C.c_void_p), ("n_verts", C.c_uint32), ("tris", C.c_void_p), ("n_tris", C.c_uint32),
                ("tri_geom", C.c_void_p), ("n_geoms", C.c_uint32), ("geom_material", C.c_void_p),
                ("geom_light", C.c_void_p), ("materials", C.c_void_p), ("n_materials", C.c_uint32),
                ("lights", C.c_void_p), ("n_lights", C.c_uint32), ("camera", OrcCamera)]


class OrcRenderParams(C.Structure):
    _fields_ = [("spp", C.c_int32), ("num_path_vertices", C.c_int32), ("seed", C.c_uint32),
                ("use_mis", C.c_int32), ("first_sample", C.c_int32), ("threads", C.c_int32),
                ("px0", C.c_int32), ("py0", C.c_int32), ("px1", C.c_int32), ("py1", C.c_int32)]


class OrcStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("extend_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("seconds", C.c_double)]


class OrcBdptRecord(C.Structure):
    _fields_ = [("s", C.c_int32), ("t", C.c_int32), ("n", C.c_int32), ("pad", C.c_int32),
                ("weight", C.c_double), ("contrib", C.c_double * 3),
                ("pos", (C.c_double * 3) * 64), ("nrm", (C.c_double * 3) * 64),
                ("kind", C.c_int32 * 64), ("id", C.c_int32 * 64),
                ("lp", C.c_double * 64), ("cp", C.c_double * 64)]


def build() -> str:
    """Compile liboracle.so (gcc); the reference harness only where the reference exists."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        L.orc_scene_create.argtypes = [C.POINTER(OrcSceneDesc), C.POINTER(P)]
        L.orc_scene_destroy.argtypes = [P]
        L.orc_last_error.restype = C.c_char_p
        L.orc_render_nee.argtypes = [P, C.POINTER(OrcRenderParams), P, P, C.POINTER(OrcStats)]
        L.orc_render_primitive.argtypes = [P, C.POINTER(OrcRenderParams), P, P, C.POINTER(OrcStats)]
        L.orc_render_visualize.argtypes = [P, C.POINTER(OrcRenderParams), P, P, C.POINTER(OrcStats)]
        L.orc_render_light.argtypes = [P, C.POINTER(OrcRenderParams), P, P, C.POINTER(OrcStats)]
        L.orc_render_bdpt.argtypes = [P, C.POINTER(OrcRenderParams), P, P, C.POINTER(OrcStats)]
        L.orc_bdpt_records.argtypes = [P, C.c_int, C.c_uint32, C.c_int, C.c_int, C.c_uint32,
                                       C.POINTER(OrcBdptRecord), C.c_int, C.POINTER(C.c_int)]
        L.orc_trace_closest.argtypes = [P, P, C.c_uint32, P, P, P]
        L.orc_trace_any.argtypes = [P, P, C.c_uint32, P]
        L.orc_trace_closest_brute.argtypes = [P, P, C.c_uint32, P, P, P, P, P]
        L.orc_trace_any_brute.argtypes = [P, P, C.c_uint32, P, P]
        L.orc_rng_u32.argtypes = [C.c_uint32] * 4
        L.orc_rng_u32.restype = C.c_uint32
        L.orc_rng_uniform.argtypes = [C.c_uint32] * 4
        L.orc_rng_uniform.restype = C.c_double
        D3 = C.POINTER(C.c_double)
        L.orc_onb.argtypes = [D3, D3, D3]
        L.orc_cosine_hemisphere.argtypes = [D3, D3]
        L.orc_barycentric_from_square.argtypes = [D3, D3]
        L.orc_disk_from_square.argtypes = [D3, D3]
        L.orc_cdf_sample.argtypes = [D3, C.c_int, C.c_double, D3, D3]
        L.orc_cdf_sample.restype = C.c_int
        L.orc_bsdf_eval.argtypes = [C.POINTER(OrcMaterial), D3, D3, D3, D3]
        L.orc_bsdf_eval_n.argtypes = [C.POINTER(OrcMaterial), C.c_int, D3, D3, D3, D3]
        L.orc_bsdf_sample.argtypes = [C.POINTER(OrcMaterial), D3, D3, D3, D3, D3]
        L.orc_light_sample.argtypes = [P, D3, D3, D3, D3, D3]
        L.orc_light_sample.restype = C.c_int
        L.orc_camera_ray.argtypes = [P, C.c_int, C.c_int, D3, D3, D3, D3, D3, D3]
        L.orc_camera_we1.argtypes = [P, D3, D3, D3]
        L.orc_camera_we1.restype = C.c_double
        L.orc_scene_set_medium.argtypes = [P, D3, D3, C.c_double]
        L.orc_scene_set_envmap.argtypes = [P, P, C.c_int, C.c_int]
        L.orc_env_eval.argtypes = [P, D3, D3, D3]
        L.orc_env_sample_ref.argtypes = [P, D3, D3, D3, D3]
        L.orc_hg_pdf.argtypes = [C.c_double, C.c_double]
        L.orc_hg_pdf.restype = C.c_double
        L.orc_hg_sample.argtypes = [C.c_double, D3, D3, D3]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _darr(x, n=3):
    return np.ascontiguousarray(np.asarray(x, np.float64).reshape(n))


def material(type_, twosided=1, rho_d=(0, 0, 0), rho_s=(0, 0, 0), sigma=0.0, exponent=0.0) -> OrcMaterial:
    m = OrcMaterial()
    m.type, m.twosided = int(type_), int(twosided)
    m.rho_d[:] = [float(v) for v in rho_d]
    m.rho_s[:] = [float(v) for v in rho_s]
    m.sigma, m.exponent = float(sigma), float(exponent)
    return m


class OracleScene:
    """An oracle scene built from wurstrenderer_amd.scenes.Scene arrays."""

    def __init__(self, sc):
        L = lib()
        self._keep = [np.ascontiguousarray(a) for a in
                      (sc.verts, sc.tris, sc.tri_geom, sc.geom_material, sc.geom_light, sc.materials, sc.lights)]
        v, t, tg, gm, gl, mats, lights = self._keep
        d = OrcSceneDesc()
        d.verts, d.n_verts = v.ctypes.data, v.shape[0]
        d.tris, d.n_tris = t.ctypes.data, t.shape[0]
        d.tri_geom, d.n_geoms = tg.ctypes.data, gm.shape[0]
        d.geom_material, d.geom_light = gm.ctypes.data, gl.ctypes.data
        d.materials, d.n_materials = mats.ctypes.data, mats.shape[0]
        d.lights, d.n_lights = lights.ctypes.data, lights.shape[0]
        C.memmove(C.byref(d.camera), sc.camera.tobytes(), C.sizeof(OrcCamera))
        h = C.c_void_p()
        if L.orc_scene_create(C.byref(d), C.byref(h)) != 0:
            raise RuntimeError("oracle: " + L.orc_last_error().decode())
        self.h = h
        self.W, self.H = sc.resolution
        if getattr(sc, "medium", None):
            self.set_medium(sc.medium["sigma_a"], sc.medium["sigma_s"], sc.medium["g"])
        if getattr(sc, "envmap", None) is not None:
            self.set_envmap(sc.envmap)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def render(self, spp, num_path_vertices=10, seed=0, use_mis=True, first_sample=0, threads=0, window=None,
               integrator="nextevent_path_tracer"):
        """Returns (base, splat, stats); image = base/spp + splat/(spp*W*H).  `integrator`
        is the reference's JSON name: nextevent_path_tracer, primitive_path_tracer,
        light_splatting, visualize or bidirectional_path_tracer."""
        fn = {"nextevent_path_tracer": lib().orc_render_nee, "primitive_path_tracer": lib().orc_render_primitive,
              "light_splatting": lib().orc_render_light, "visualize": lib().orc_render_visualize,
              "bidirectional_path_tracer": lib().orc_render_bdpt}[integrator]
        base = np.zeros((self.H, self.W, 3), np.float64)
        splat = np.zeros((self.H, self.W, 3), np.float64)
        p = OrcRenderParams(spp, num_path_vertices, seed, int(bool(use_mis)), first_sample, threads, 0, 0, 0, 0)
        if window is not None:
            p.px0, p.py0, p.px1, p.py1 = window
        st = OrcStats()
        if fn(self.h, C.byref(p), base.ctypes.data, splat.ctypes.data, C.byref(st)) != 0:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())
        return base, splat, st

    def bdpt_records(self, num_path_vertices, seed, px, py, sample, max_records=4096):
        """The strategies of one bidirectional camera sample (orc_bdpt_records) as dicts
        with s, t, n, weight, contrib, pos (n, 3), nrm (n, 3), kind, id, lp, cp."""
        buf = (OrcBdptRecord * max_records)()
        n = C.c_int()
        if lib().orc_bdpt_records(self.h, num_path_vertices, seed, px, py, sample, buf, max_records, C.byref(n)) != 0:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())
        out = []
        for r in buf[:n.value]:
            k = r.n
            out.append(dict(s=r.s, t=r.t, n=k, weight=r.weight, contrib=np.array(r.contrib[:]),
                            pos=np.array([r.pos[i][:] for i in range(k)]), nrm=np.array([r.nrm[i][:] for i in range(k)]),
                            kind=np.array(r.kind[:k]), id=np.array(r.id[:k]),
                            lp=np.array(r.lp[:k]), cp=np.array(r.cp[:k])))
        return out

    def image(self, spp, **kw):
        base, splat, st = self.render(spp, **kw)
        return base / spp + splat / (spp * self.W * self.H), st

    def trace_closest(self, rays, brute=False):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        geom = np.zeros(n, np.uint32)
        prim = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        if brute:
            sec = np.zeros(n, np.float32)
            mar = np.zeros(n, np.float32)
            lib().orc_trace_closest_brute(self.h, rays.ctypes.data, n, geom.ctypes.data, prim.ctypes.data,
                                          t.ctypes.data, sec.ctypes.data, mar.ctypes.data)
            return geom, prim, t, sec, mar
        lib().orc_trace_closest(self.h, rays.ctypes.data, n, geom.ctypes.data, prim.ctypes.data, t.ctypes.data)
        return geom, prim, t

    def trace_any(self, rays, brute=False):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        occ = np.zeros(n, np.uint8)
        if brute:
            amb = np.zeros(n, np.uint8)
            lib().orc_trace_any_brute(self.h, rays.ctypes.data, n, occ.ctypes.data, amb.ctypes.data)
            return occ, amb
        lib().orc_trace_any(self.h, rays.ctypes.data, n, occ.ctypes.data)
        return occ

    def set_envmap(self, rgb):
        """Equirectangular envmap light, float (H, 2H, 3); None removes it."""
        if rgb is None:
            lib().orc_scene_set_envmap(self.h, None, 0, 0)
            return
        a = np.ascontiguousarray(rgb, np.float32)
        self._env = a
        if lib().orc_scene_set_envmap(self.h, a.ctypes.data, a.shape[1], a.shape[0]) != 0:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())

    def env_eval(self, d):
        """EquiRectRep eval (bilinear) and pdfA at direction d."""
        dd = np.ascontiguousarray(d, np.float64)
        out = np.zeros(3)
        pdf = np.zeros(1)
        lib().orc_env_eval(self.h, _dp(dd), _dp(out), _dp(pdf))
        return out, float(pdf[0])

    def env_sample_ref(self, u):
        """EquiRectRep::sample with the reference's use of the two numbers: (dir, pdf, value)."""
        uu = np.ascontiguousarray(u, np.float64)
        d, pdf, v = np.zeros(3), np.zeros(1), np.zeros(3)
        lib().orc_env_sample_ref(self.h, _dp(uu), _dp(d), _dp(pdf), _dp(v))
        return d, float(pdf[0]), v

    def set_medium(self, sigma_a=(0, 0, 0), sigma_s=(0, 0, 0), g=0.0):
        """Homogeneous medium filling the scene, HG phase (zeros remove it)."""
        a, s = _darr(sigma_a), _darr(sigma_s)
        if lib().orc_scene_set_medium(self.h, _dp(a), _dp(s), float(g)) != 0:
            raise RuntimeError(lib().orc_last_error().decode())

    def light_sample(self, u):
        u = _darr(u, 2)
        pos, nrm, con, pdf = np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(1)
        li = lib().orc_light_sample(self.h, _dp(u), _dp(pos), _dp(nrm), _dp(con), _dp(pdf))
        return li, pos, nrm, con, float(pdf[0])

    def camera_ray(self, px, py, ulens, upix):
        ul, up = _darr(ulens, 2), _darr(upix, 2)
        o, d, r, pdf = np.zeros(3), np.zeros(3), np.zeros(2), np.zeros(1)
        lib().orc_camera_ray(self.h, int(px), int(py), _dp(ul), _dp(up), _dp(o), _dp(d), _dp(r), _dp(pdf))
        return o, d, r, float(pdf[0])

    def camera_we1(self, cam_pos, w):
        cp, ww = _darr(cam_pos), _darr(w)
        r = np.full(2, -1.0)
        v = lib().orc_camera_we1(self.h, _dp(cp), _dp(ww), _dp(r))
        return v, r


def onb(n):
    n = _darr(n)
    x, z = np.zeros(3), np.zeros(3)
    lib().orc_onb(_dp(n), _dp(x), _dp(z))
    return x, z


def cosine_hemisphere(u):
    u, o = _darr(u, 2), np.zeros(3)
    lib().orc_cosine_hemisphere(_dp(u), _dp(o))
    return o


def barycentric_from_square(u):
    u, o = _darr(u, 2), np.zeros(2)
    lib().orc_barycentric_from_square(_dp(u), _dp(o))
    return o


def disk_from_square(u):
    u, o = _darr(u, 2), np.zeros(2)
    lib().orc_disk_from_square(_dp(u), _dp(o))
    return o


def cdf_sample(weights, u):
    w = np.ascontiguousarray(np.asarray(weights, np.float64))
    pdf, rem = np.zeros(1), np.zeros(1)
    i = lib().orc_cdf_sample(_dp(w), len(w), float(u), _dp(pdf), _dp(rem))
    return i, float(pdf[0]), float(rem[0])


def bsdf_eval(m: OrcMaterial, wi, wo):
    wi, wo, f, pdf = _darr(wi), _darr(wo), np.zeros(3), np.zeros(1)
    lib().orc_bsdf_eval(C.byref(m), _dp(wi), _dp(wo), _dp(f), _dp(pdf))
    return f, float(pdf[0])


def bsdf_eval_n(m: OrcMaterial, wi, wo):
    wi = np.ascontiguousarray(np.broadcast_to(np.asarray(wi, np.float64), np.shape(wo)))
    wo = np.ascontiguousarray(np.asarray(wo, np.float64))
    n = wo.shape[0]
    f, pdf = np.zeros((n, 3)), np.zeros(n)
    lib().orc_bsdf_eval_n(C.byref(m), n, _dp(wi), _dp(wo), _dp(f), _dp(pdf))
    return f, pdf


def bsdf_sample(m: OrcMaterial, wi, u):
    wi, u = _darr(wi), _darr(u, 2)
    wo, w, pdf = np.zeros(3), np.zeros(3), np.zeros(1)
    lib().orc_bsdf_sample(C.byref(m), _dp(wi), _dp(u), _dp(wo), _dp(w), _dp(pdf))
    return wo, w, float(pdf[0])


def hg_pdf(g, cos_theta) -> float:
    return lib().orc_hg_pdf(float(g), float(cos_theta))


def hg_sample(g, fwd, u):
    f, uu, out = _darr(fwd), _darr(u, 2), np.zeros(3)
    lib().orc_hg_sample(float(g), _dp(f), _dp(uu), _dp(out))
    return out


def rng_uniform(seed, pixel, sample, dim) -> float:
    return lib().orc_rng_uniform(seed, pixel, sample, dim)


def rng_u32(seed, pixel, sample, dim) -> int:
    return lib().orc_rng_u32(seed, pixel, sample, dim)


def rel_mse(img, ref, eps=1e-2) -> float:
    """mean over pixels and channels of (I-R)^2/(R^2+eps)."""
    img = np.asarray(img, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + eps)))
