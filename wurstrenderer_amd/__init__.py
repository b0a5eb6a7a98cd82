"""MI355X-native `nextevent_path_tracer`: Python binding of libwurst_amd.so.

The library exposes the C-ABI declared in include/wurst_amd.h.  This module is
the ctypes stub over that ABI (the same binding a ctypes user of the reference
path would write); the compute runs in the HIP kernels of the library.  There
is no CPU fallback: if the library is missing or no GPU is present the calls
raise.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import subprocess
import weakref

import numpy as np

from . import scenes  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwurst_amd.so")

WR_OK = 0
WR_INVALID_ID = 0xFFFFFFFF


class WrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"wurst_amd error {code}: {msg}")
        self.code = code


class _Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("lookat", C.c_float * 3), ("up", C.c_float * 3),
                ("fovy_deg", C.c_float), ("width", C.c_int32), ("height", C.c_int32),
                ("aperture_radius", C.c_float), ("focal_distance", C.c_float)]


class _SceneDesc(C.Structure):
    _fields_ = [("verts", C.c_void_p), ("n_verts", C.c_uint32), ("tris", C.c_void_p), ("n_tris", C.c_uint32),
                ("tri_geom", C.c_void_p), ("n_geoms", C.c_uint32), ("geom_material", C.c_void_p),
                ("geom_light", C.c_void_p), ("materials", C.c_void_p), ("n_materials", C.c_uint32),
                ("lights", C.c_void_p), ("n_lights", C.c_uint32), ("camera", _Camera)]


class RenderParams(C.Structure):
    _fields_ = [("spp", C.c_int32), ("num_path_vertices", C.c_int32), ("seed", C.c_uint32),
                ("use_mis", C.c_int32), ("first_sample", C.c_int32), ("max_paths_in_flight", C.c_int32),
                ("profile", C.c_int32), ("count_traversal", C.c_int32), ("use_graph", C.c_int32),
                ("fetch_chunk", C.c_uint32), ("refill_lanes", C.c_uint32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchJob(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("film", C.c_void_p), ("spp", C.c_int32), ("num_path_vertices", C.c_int32),
                ("seed", C.c_uint32), ("use_mis", C.c_int32), ("first_sample", C.c_int32)]


class RenderStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("extend_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("wall_ms", C.c_double), ("extend_ms", C.c_double), ("shadow_ms", C.c_double),
                ("shade_ms", C.c_double), ("generate_ms", C.c_double), ("accumulate_ms", C.c_double),
                ("extend_launches", C.c_uint64), ("shadow_launches", C.c_uint64),
                ("extend_nodes", C.c_uint64), ("extend_tris", C.c_uint64),
                ("shadow_nodes", C.c_uint64), ("shadow_tris", C.c_uint64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Medium(C.Structure):
    _fields_ = [("sigma_a", C.c_float * 3), ("sigma_s", C.c_float * 3), ("g", C.c_float)]


class GuideConfig(C.Structure):
    _fields_ = [("learning_rate", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("epsilon", C.c_float),
                ("batch_size", C.c_uint32), ("record_capacity", C.c_uint32), ("seed", C.c_uint32),
                ("avg_decay", C.c_float)]


class GuidedParams(C.Structure):
    _fields_ = [("guide_fraction", C.c_float), ("spp_per_pass", C.c_uint32), ("train_steps", C.c_uint32),
                ("warmup_passes", C.c_uint32), ("fp8_inference", C.c_int32), ("max_guided_bounce", C.c_uint32),
                ("overlap_training", C.c_uint32)]


class GuideStats(C.Structure):
    _fields_ = [("records", C.c_uint64), ("train_steps", C.c_uint64), ("train_ms", C.c_double),
                ("last_loss", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


GUIDE_NUM_PARAMS = 17192
GUIDE_NUM_OUTPUTS = 40

BVH_GPU, BVH_HOST = 0, 1
BUILDERS = {"gpu": BVH_GPU, "host": BVH_HOST}


class BuildParams(C.Structure):
    _fields_ = [("builder", C.c_int32), ("ploc_radius", C.c_int32)]


class BvhStats(C.Structure):
    _fields_ = [("builder", C.c_int32), ("depth", C.c_uint32), ("max_stack", C.c_uint32), ("iterations", C.c_uint32),
                ("n_nodes", C.c_uint64), ("build_ms", C.c_double), ("sah_cost", C.c_double), ("hash", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# name -> (restype, argtypes); every symbol declared in include/wurst_amd.h
P = C.c_void_p
EXPORTS = {
    "wr_last_error": (C.c_char_p, []),
    "wr_version": (C.c_char_p, []),
    "wr_device_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "wr_device_destroy": (None, [P]),
    "wr_scene_validate": (C.c_int, [C.POINTER(_SceneDesc)]),
    "wr_bvh_selftest": (C.c_int, [C.POINTER(_SceneDesc), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32),
                                  C.POINTER(C.c_uint32)]),
    "wr_scene_create": (C.c_int, [P, C.POINTER(_SceneDesc), C.POINTER(P)]),
    "wr_scene_create_ex": (C.c_int, [P, C.POINTER(_SceneDesc), C.POINTER(BuildParams), C.POINTER(P)]),
    "wr_scene_bvh_check": (C.c_int, [P, C.POINTER(_SceneDesc), C.POINTER(BvhStats)]),
    "wr_scene_destroy": (None, [P]),
    "wr_scene_info": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32),
                                C.POINTER(C.c_uint64)]),
    "wr_scene_set_medium": (C.c_int, [P, C.POINTER(Medium)]),
    "wr_scene_set_envmap": (C.c_int, [P, P, C.c_uint32, C.c_uint32]),
    "wr_film_create": (C.c_int, [P, C.POINTER(P)]),
    "wr_film_destroy": (None, [P]),
    "wr_film_clear": (C.c_int, [P]),
    "wr_film_read": (C.c_int, [P, P, P]),
    "wr_render_params_validate": (C.c_int, [C.POINTER(RenderParams)]),
    "wr_render_nee": (C.c_int, [P, P, C.POINTER(RenderParams), C.POINTER(RenderStats)]),
    "wr_render_nee_noalloc": (C.c_int, [P, P, C.POINTER(RenderParams), C.POINTER(RenderStats)]),
    "wr_render_primitive": (C.c_int, [P, P, C.POINTER(RenderParams), C.POINTER(RenderStats)]),
    "wr_render_visualize": (C.c_int, [P, P, C.POINTER(RenderParams), C.POINTER(RenderStats)]),
    "wr_render_light": (C.c_int, [P, P, C.POINTER(RenderParams), C.POINTER(RenderStats)]),
    "wr_render_bdpt": (C.c_int, [P, P, C.POINTER(RenderParams), C.POINTER(RenderStats)]),
    "wr_trace_closest": (C.c_int, [P, P, C.c_uint32, P, P, P]),
    "wr_trace_any": (C.c_int, [P, P, C.c_uint32, P]),
    "wr_trace_closest_device": (C.c_int, [P, P, C.c_uint32, P, C.POINTER(C.c_float)]),
    "wr_trace_any_device": (C.c_int, [P, P, C.c_uint32, P, C.POINTER(C.c_float)]),
    "wr_debug_bsdf_eval": (C.c_int, [P, P, P, P, C.c_uint32, P, P]),
    "wr_debug_bsdf_sample": (C.c_int, [P, P, P, P, C.c_uint32, P, P, P]),
    "wr_debug_light_sample": (C.c_int, [P, P, C.c_uint32, P, P, P, P]),
    "wr_debug_camera_ray": (C.c_int, [P, P, P, P, C.c_uint32, P, P, P]),
    "wr_debug_camera_we1": (C.c_int, [P, P, P, C.c_uint32, P, P]),
    "wr_debug_l1_request_probe": (C.c_int, [P, C.c_uint64, C.c_uint32, P, P, P]),
    "wr_render_batch": (C.c_int, [C.POINTER(BatchJob), C.c_uint32, C.POINTER(RenderParams), C.POINTER(RenderStats)]),
    "wr_autotune": (C.c_int, [C.POINTER(BatchJob), C.c_uint32, C.POINTER(RenderParams), C.POINTER(RenderParams),
                              C.POINTER(C.c_double)]),
    "wr_guide_config_default": (None, [C.POINTER(GuideConfig)]),
    "wr_guide_create": (C.c_int, [P, C.POINTER(GuideConfig), C.POINTER(P)]),
    "wr_guide_destroy": (None, [P]),
    "wr_guide_get_params": (C.c_int, [P, P]),
    "wr_guide_set_params": (C.c_int, [P, P]),
    "wr_guide_forward": (C.c_int, [P, P, P, P, C.c_uint32, P]),
    "wr_guide_forward_fp8": (C.c_int, [P, P, P, P, C.c_uint32, P]),
    "wr_guide_gradient": (C.c_int, [P, P, P, P, P, P, C.c_uint32, P, C.POINTER(C.c_double)]),
    "wr_guide_train": (C.c_int, [P, P, P, P, P, P, C.c_uint32, C.c_uint32, C.POINTER(C.c_double)]),
    "wr_guide_mixture": (C.c_int, [P, P, P, P, C.c_uint32, P, P, P]),
    "wr_render_guided": (C.c_int, [P, P, P, C.POINTER(RenderParams), C.POINTER(GuidedParams), C.POINTER(RenderStats),
                                   C.POINTER(GuideStats)]),
}

_lib = None


def build() -> str:
    """Compile libwurst_amd.so in-tree (hipcc, gfx950)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib():
    """Load libwurst_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C wurstrenderer_amd")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


# Live handles, released at interpreter exit in dependency order (guides and
# films before their scene, scenes before their device) while the HIP runtime
# is still intact, rather than in whatever order finalisation reaches them.
_LIVE = {k: weakref.WeakSet() for k in ("guide", "film", "scene", "device")}


@atexit.register
def _release_all():
    for kind in ("guide", "film", "scene", "device"):
        for obj in list(_LIVE[kind]):
            obj.close()


def _check(rc: int):
    if rc != WR_OK:
        raise WrError(rc, lib().wr_last_error().decode())


class Device:
    def __init__(self, ordinal: int = 0):
        h = C.c_void_p()
        _check(lib().wr_device_create(ordinal, C.byref(h)))
        self.h = h
        _LIVE["device"].add(self)

    def close(self):
        """Releases the handle (idempotent)."""
        if getattr(self, "h", None):
            lib().wr_device_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def scene_desc(sc: scenes.Scene):
    """Builds the C descriptor; returns (desc, keepalive arrays)."""
    keep = [np.ascontiguousarray(a) for a in
            (sc.verts, sc.tris, sc.tri_geom, sc.geom_material, sc.geom_light, sc.materials, sc.lights)]
    v, t, tg, gm, gl, mats, lights = keep
    d = _SceneDesc()
    d.verts, d.n_verts = v.ctypes.data, v.shape[0]
    d.tris, d.n_tris = t.ctypes.data, t.shape[0]
    d.tri_geom, d.n_geoms = tg.ctypes.data, gm.shape[0]
    d.geom_material, d.geom_light = gm.ctypes.data, gl.ctypes.data
    d.materials, d.n_materials = mats.ctypes.data, mats.shape[0]
    d.lights, d.n_lights = lights.ctypes.data, lights.shape[0]
    C.memmove(C.byref(d.camera), sc.camera.tobytes(), C.sizeof(_Camera))
    return d, keep


def validate(sc: scenes.Scene) -> None:
    """Host-side scene checks (raises WrError)."""
    d, keep = scene_desc(sc)
    _check(lib().wr_scene_validate(C.byref(d)))


def validate_params(params: RenderParams) -> None:
    """Host-side render-parameter checks (raises WrError)."""
    _check(lib().wr_render_params_validate(C.byref(params)))


def debug_bsdf_eval(device: Device, mats, wi, wo):
    """wr_debug_bsdf_eval: the device BSDF's f and pdf per record (local frame, y = normal);
    mats is an array of scenes.MATERIAL_DTYPE, one per record."""
    m = np.ascontiguousarray(mats, scenes.MATERIAL_DTYPE)
    a = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
    b = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
    n = len(m)
    assert len(a) == n and len(b) == n
    f, pdf = np.zeros((n, 3), np.float32), np.zeros(n, np.float32)
    _check(lib().wr_debug_bsdf_eval(device.h, m.ctypes.data, a.ctypes.data, b.ctypes.data, n, f.ctypes.data,
                                    pdf.ctypes.data))
    return f, pdf


def l1_request_probe(device: Device, footprint: int = 1 << 22, loads_per_lane: int = 4096):
    """wr_debug_l1_request_probe: one timed launch in which every lane of a full
    grid issues loads_per_lane scattered 16-byte loads (eight in flight) into a
    buffer of `footprint` bytes; returns (ms, loads, compute units)."""
    ms, loads, ncu = C.c_double(), C.c_uint64(), C.c_uint32()
    _check(lib().wr_debug_l1_request_probe(device.h, int(footprint), int(loads_per_lane), C.byref(ms), C.byref(loads),
                                           C.byref(ncu)))
    return ms.value, loads.value, ncu.value


def debug_bsdf_sample(device: Device, mats, wi, u):
    """wr_debug_bsdf_sample: the device BSDF's sampled direction, f cos / pdf and pdf per record."""
    m = np.ascontiguousarray(mats, scenes.MATERIAL_DTYPE)
    a = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
    uu = np.ascontiguousarray(u, np.float32).reshape(-1, 2)
    n = len(m)
    assert len(a) == n and len(uu) == n
    wo, w, pdf = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros(n, np.float32)
    _check(lib().wr_debug_bsdf_sample(device.h, m.ctypes.data, a.ctypes.data, uu.ctypes.data, n, wo.ctypes.data,
                                      w.ctypes.data, pdf.ctypes.data))
    return wo, w, pdf


def bvh_selftest(sc: scenes.Scene) -> dict:
    """Builds the BVH on the host and verifies its invariants (raises WrError)."""
    d, keep = scene_desc(sc)
    nn, ms, dp = C.c_uint64(), C.c_uint32(), C.c_uint32()
    _check(lib().wr_bvh_selftest(C.byref(d), C.byref(nn), C.byref(ms), C.byref(dp)))
    return dict(nodes=nn.value, max_stack=ms.value, depth=dp.value)


class GpuScene:
    """A scene resident in HBM (BVH, triangles, materials, lights, camera)."""

    def __init__(self, device: Device, sc: scenes.Scene, builder: str = "gpu", ploc_radius: int = 0):
        """builder: "gpu" (PLOC on the GPU, the default of wr_scene_create) or "host"
        (binned SAH); ploc_radius: the GPU builder's search window (0 = default)."""
        if builder not in BUILDERS:
            raise ValueError(f"builder must be one of {sorted(BUILDERS)}")
        d, keep = scene_desc(sc)
        h = C.c_void_p()
        bp = BuildParams(BUILDERS[builder], int(ploc_radius))
        _check(lib().wr_scene_create_ex(device.h, C.byref(d), C.byref(bp), C.byref(h)))
        self.h = h
        _LIVE["scene"].add(self)
        self.device = device
        self.W, self.H = sc.resolution
        self.builder = builder
        del keep
        if getattr(sc, "medium", None):
            self.set_medium(sc.medium["sigma_a"], sc.medium["sigma_s"], sc.medium["g"])
        if getattr(sc, "envmap", None) is not None:
            self.set_envmap(sc.envmap)

    def set_envmap(self, rgb):
        """Equirectangular envmap light: float (H, 2H, 3); None removes it."""
        if rgb is None:
            _check(lib().wr_scene_set_envmap(self.h, None, 0, 0))
            return
        a = np.ascontiguousarray(rgb, np.float32)
        if a.ndim != 3 or a.shape[2] != 3:
            raise ValueError(f"envmap must be (H, W, 3), got {a.shape}")
        _check(lib().wr_scene_set_envmap(self.h, a.ctypes.data, a.shape[1], a.shape[0]))

    def set_medium(self, sigma_a=(0, 0, 0), sigma_s=(0, 0, 0), g=0.0):
        """Homogeneous medium filling the scene with an HG phase function (zeros: vacuum)."""
        m = Medium()
        m.sigma_a[:] = [float(x) for x in sigma_a]
        m.sigma_s[:] = [float(x) for x in sigma_s]
        m.g = float(g)
        _check(lib().wr_scene_set_medium(self.h, C.byref(m)))

    def close(self):
        """Releases the handle (idempotent), after the guides and films of this
        scene, which the library requires to go first."""
        if getattr(self, "h", None):
            for kind in ("guide", "film"):
                for obj in list(_LIVE[kind]):
                    if getattr(obj, "scene", None) is self:
                        obj.close()
            lib().wr_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def info(self) -> dict:
        nn, nt, nb = C.c_uint64(), C.c_uint64(), C.c_uint64()
        ms = C.c_uint32()
        _check(lib().wr_scene_info(self.h, C.byref(nn), C.byref(nt), C.byref(ms), C.byref(nb)))
        return dict(nodes=nn.value, tris=nt.value, max_stack=ms.value, device_bytes=nb.value)

    def bvh_check(self, sc: scenes.Scene) -> dict:
        """Downloads the BVH, checks its invariants against `sc` (the scene this
        was built from; raises WrError on a violation) and returns its statistics."""
        d, keep = scene_desc(sc)
        st = BvhStats()
        _check(lib().wr_scene_bvh_check(self.h, C.byref(d), C.byref(st)))
        del keep
        return st.as_dict()

    def debug_light_sample(self, u):
        """wr_debug_light_sample: the device's light sampling at uniforms u (n, 2)."""
        u = np.ascontiguousarray(u, np.float32).reshape(-1, 2)
        n = len(u)
        pos, nrm, con, pdf = (np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32),
                              np.zeros((n, 3), np.float32), np.zeros(n, np.float32))
        _check(lib().wr_debug_light_sample(self.h, u.ctypes.data, n, pos.ctypes.data, nrm.ctypes.data,
                                           con.ctypes.data, pdf.ctypes.data))
        return pos, nrm, con, pdf

    def debug_camera_ray(self, pixel, ulens, upix):
        """wr_debug_camera_ray: lens point, direction and per-pixel pdfW of camera samples."""
        px = np.ascontiguousarray(pixel, np.uint32).reshape(-1, 2)
        ul = np.ascontiguousarray(ulens, np.float32).reshape(-1, 2)
        up = np.ascontiguousarray(upix, np.float32).reshape(-1, 2)
        n = len(px)
        org, d, pdf = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros(n, np.float32)
        _check(lib().wr_debug_camera_ray(self.h, px.ctypes.data, ul.ctypes.data, up.ctypes.data, n, org.ctypes.data,
                                         d.ctypes.data, pdf.ctypes.data))
        return org, d, pdf

    def debug_camera_we1(self, org, w):
        """wr_debug_camera_we1: We1 and raster position toward directions w from lens points org."""
        o = np.ascontiguousarray(org, np.float32).reshape(-1, 3)
        ww = np.ascontiguousarray(w, np.float32).reshape(-1, 3)
        n = len(o)
        we1, r = np.zeros(n, np.float32), np.zeros((n, 2), np.float32)
        _check(lib().wr_debug_camera_we1(self.h, o.ctypes.data, ww.ctypes.data, n, we1.ctypes.data, r.ctypes.data))
        return we1, r

    def trace_closest(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        geom = np.zeros(n, np.uint32)
        prim = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        _check(lib().wr_trace_closest(self.h, rays.ctypes.data, n, geom.ctypes.data, prim.ctypes.data, t.ctypes.data))
        return geom, prim, t

    def trace_any(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        occ = np.zeros(n, np.uint8)
        _check(lib().wr_trace_any(self.h, rays.ctypes.data, n, occ.ctypes.data))
        return occ


class Film:
    """FP64 base + splat buffers in HBM (Film, src/common/film.h:35-68)."""

    def __init__(self, scene: GpuScene):
        h = C.c_void_p()
        _check(lib().wr_film_create(scene.h, C.byref(h)))
        self.h = h
        _LIVE["film"].add(self)
        self.scene = scene
        self.spp = 0

    def close(self):
        """Releases the handle (idempotent)."""
        if getattr(self, "h", None):
            lib().wr_film_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def clear(self):
        _check(lib().wr_film_clear(self.h))
        self.spp = 0

    def read(self):
        W, H = self.scene.W, self.scene.H
        base = np.zeros((H, W, 3), np.float64)
        splat = np.zeros((H, W, 3), np.float64)
        _check(lib().wr_film_read(self.h, base.ctypes.data, splat.ctypes.data))
        return base, splat

    def image(self):
        """Film::getImage: base/spp + splat/(spp*W*H) (film.h:78-90, pathnextevent.h:154-157)."""
        base, splat = self.read()
        W, H = self.scene.W, self.scene.H
        return base / self.spp + splat / (self.spp * W * H)


def render_params(spp=0, num_path_vertices=10, seed=0, use_mis=True, first_sample=0, max_paths_in_flight=0,
                  profile=False, count_traversal=False, use_graph=False, fetch_chunk=0, refill_lanes=0,
                  tuning: RenderParams | None = None) -> RenderParams:
    """wr_render_params; `tuning` (from autotune) supplies the scheduling fields."""
    p = RenderParams(spp, num_path_vertices, seed, int(bool(use_mis)), first_sample, max_paths_in_flight,
                     int(bool(profile)), int(bool(count_traversal)), int(bool(use_graph)), fetch_chunk, refill_lanes)
    if tuning is not None:
        for k in ("max_paths_in_flight", "use_graph", "fetch_chunk", "refill_lanes"):
            setattr(p, k, getattr(tuning, k))
    return p


def render_nee(scene: GpuScene, film: Film, spp: int, num_path_vertices: int = 10, seed: int = 0,
               use_mis: bool = True, first_sample: int | None = None, max_paths_in_flight: int = 0,
               profile: bool = False, count_traversal: bool = False, use_graph: bool = False,
               fetch_chunk: int = 0, refill_lanes: int = 0, tuning: RenderParams | None = None) -> RenderStats:
    """PathNeeRenderer::render on the GPU: adds `spp` samples per pixel into `film`."""
    if first_sample is None:
        first_sample = film.spp
    p = render_params(spp, num_path_vertices, seed, use_mis, first_sample, max_paths_in_flight, profile,
                      count_traversal, use_graph, fetch_chunk, refill_lanes, tuning)
    st = RenderStats()
    _check(lib().wr_render_nee(scene.h, film.h, C.byref(p), C.byref(st)))
    film.spp += spp
    return st


def render_primitive(scene: GpuScene, film: Film, spp: int, num_path_vertices: int = 10, seed: int = 0,
                     first_sample: int | None = None, max_paths_in_flight: int = 0, profile: bool = False,
                     count_traversal: bool = False) -> RenderStats:
    """PrimitivePathRenderer::render on the GPU (primitivepath.h:42-96): BSDF-sampled
    paths that count the emitters they hit, without NEE; adds `spp` samples per pixel."""
    if first_sample is None:
        first_sample = film.spp
    p = render_params(spp, num_path_vertices, seed, False, first_sample, max_paths_in_flight, profile, count_traversal)
    st = RenderStats()
    _check(lib().wr_render_primitive(scene.h, film.h, C.byref(p), C.byref(st)))
    film.spp += spp
    return st


def render_light(scene: GpuScene, film: Film, spp: int, num_path_vertices: int = 10, seed: int = 0,
                 first_sample: int | None = None, max_paths_in_flight: int = 0, profile: bool = False,
                 count_traversal: bool = False) -> RenderStats:
    """LightSplattingRenderer::render on the GPU (lightsplatting.h:37-113): spp*W*H light
    subpaths whose vertices connect to the camera; film.image() = splat/(spp*W*H)."""
    if first_sample is None:
        first_sample = film.spp
    p = render_params(spp, num_path_vertices, seed, False, first_sample, max_paths_in_flight, profile, count_traversal)
    st = RenderStats()
    _check(lib().wr_render_light(scene.h, film.h, C.byref(p), C.byref(st)))
    film.spp += spp
    return st


def render_bdpt(scene: GpuScene, film: Film, spp: int, num_path_vertices: int = 10, seed: int = 0,
                first_sample: int | None = None, profile: bool = False) -> RenderStats:
    """BidirPathRenderer::render on the GPU (bidirpath.h:316-395): per camera sample a
    light and a camera subpath and every strategy s + t <= N, balance-heuristic
    weights; film.image() = base/spp + splat/(spp*W*H) as for NEE."""
    if first_sample is None:
        first_sample = film.spp
    p = render_params(spp, num_path_vertices, seed, True, first_sample, 0, profile)
    st = RenderStats()
    _check(lib().wr_render_bdpt(scene.h, film.h, C.byref(p), C.byref(st)))
    film.spp += spp
    return st


def render_visualize(scene: GpuScene, film: Film, spp: int, seed: int = 0, first_sample: int | None = None,
                     profile: bool = False) -> RenderStats:
    """VisualizeGeometryNormal::render on the GPU (visualizegeomnormal.h:32-74): the
    primary hit's unit geometric normal as RGB, (1, 1, 1) on emitters, averaged over spp."""
    if first_sample is None:
        first_sample = film.spp
    p = render_params(spp, 2, seed, False, first_sample, 0, profile)
    st = RenderStats()
    _check(lib().wr_render_visualize(scene.h, film.h, C.byref(p), C.byref(st)))
    film.spp += spp
    return st


def _f32(a, cols=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if cols is not None and (a.ndim != 2 or a.shape[1] != cols):
        raise ValueError(f"expected an (n, {cols}) array, got {a.shape}")
    return a


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Guide:
    """The path-guiding network bound to a scene (wr_guide_* of the C-ABI)."""

    def __init__(self, scene: GpuScene, learning_rate: float = 1e-2, batch_size: int = 1 << 18,
                 record_capacity: int = 1 << 20, seed: int = 0, avg_decay: float | None = None):
        """avg_decay: decay of the parameter average the renderer's inference uses
        (None: the library default, 0.99; 0: the parameters themselves)."""
        cfg = GuideConfig()
        lib().wr_guide_config_default(C.byref(cfg))
        cfg.learning_rate = learning_rate
        cfg.batch_size = batch_size
        cfg.record_capacity = record_capacity
        cfg.seed = seed
        if avg_decay is not None:
            cfg.avg_decay = avg_decay
        h = C.c_void_p()
        _check(lib().wr_guide_create(scene.h, C.byref(cfg), C.byref(h)))
        self.h = h
        _LIVE["guide"].add(self)
        self.scene = scene

    def close(self):
        """Releases the handle (idempotent)."""
        if getattr(self, "h", None):
            lib().wr_guide_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def get_params(self) -> np.ndarray:
        p = np.empty(GUIDE_NUM_PARAMS, dtype=np.float32)
        _check(lib().wr_guide_get_params(self.h, _ptr(p)))
        return p

    def set_params(self, params) -> None:
        p = _f32(params).ravel()
        if p.size != GUIDE_NUM_PARAMS:
            raise ValueError("wrong parameter count")
        _check(lib().wr_guide_set_params(self.h, _ptr(p)))

    def forward(self, pos, nrm, wo, fp8: bool = False) -> np.ndarray:
        """Raw outputs (n, 40); fp8 selects the e4m3 inference variant."""
        pos, nrm, wo = _f32(pos, 3), _f32(nrm, 3), _f32(wo, 3)
        out = np.empty((len(pos), GUIDE_NUM_OUTPUTS), dtype=np.float32)
        fn = lib().wr_guide_forward_fp8 if fp8 else lib().wr_guide_forward
        _check(fn(self.h, _ptr(pos), _ptr(nrm), _ptr(wo), len(pos), _ptr(out)))
        return out

    def gradient(self, pos, nrm, wo, wi, weight):
        pos, nrm, wo, wi = _f32(pos, 3), _f32(nrm, 3), _f32(wo, 3), _f32(wi, 3)
        w = _f32(weight).ravel()
        g = np.empty(GUIDE_NUM_PARAMS, dtype=np.float32)
        loss = C.c_double()
        _check(lib().wr_guide_gradient(self.h, _ptr(pos), _ptr(nrm), _ptr(wo), _ptr(wi), _ptr(w), len(pos), _ptr(g),
                                       C.byref(loss)))
        return loss.value, g

    def train(self, pos, nrm, wo, wi, weight, steps: int) -> float:
        pos, nrm, wo, wi = _f32(pos, 3), _f32(nrm, 3), _f32(wo, 3), _f32(wi, 3)
        w = _f32(weight).ravel()
        loss = C.c_double()
        _check(lib().wr_guide_train(self.h, _ptr(pos), _ptr(nrm), _ptr(wo), _ptr(wi), _ptr(w), len(pos), steps,
                                    C.byref(loss)))
        return loss.value

    def mixture(self, raw, dirs, u):
        raw, dirs, u = _f32(raw, GUIDE_NUM_OUTPUTS), _f32(dirs, 3), _f32(u, 3)
        n = len(raw)
        pd = np.empty(n, dtype=np.float32)
        smp = np.empty((n, 3), dtype=np.float32)
        sp = np.empty(n, dtype=np.float32)
        _check(lib().wr_guide_mixture(self.h, _ptr(raw), _ptr(dirs), _ptr(u), n, _ptr(pd), _ptr(smp), _ptr(sp)))
        return pd, smp, sp


def render_guided(scene: GpuScene, film: Film, guide: Guide, spp: int, num_path_vertices: int = 10, seed: int = 0,
                  guide_fraction: float = 0.5, spp_per_pass: int = 4, train_steps: int = 4, warmup_passes: int = 1,
                  use_mis: bool = True, first_sample: int | None = None, max_paths_in_flight: int = 0,
                  profile: bool = False, fp8_inference: bool = False, use_graph: bool = False,
                  max_guided_bounce: int = 0, overlap_training: bool = False):
    """wr_render_guided: progressive passes of spp_per_pass, training between passes;
    fp8_inference runs the renderer's network inference in e4m3; use_graph replays
    each pass's iterations as a captured hipGraph; overlap_training renders pass q
    with the network trained on pass q - 2 while the training on pass q - 1 runs."""
    if first_sample is None:
        first_sample = film.spp
    p = RenderParams(spp, num_path_vertices, seed, int(bool(use_mis)), first_sample, max_paths_in_flight,
                     int(bool(profile)), 0, int(bool(use_graph)))
    gp = GuidedParams(guide_fraction, spp_per_pass, train_steps, warmup_passes, int(bool(fp8_inference)),
                      max_guided_bounce, int(bool(overlap_training)))
    st, gs = RenderStats(), GuideStats()
    _check(lib().wr_render_guided(scene.h, film.h, guide.h, C.byref(p), C.byref(gp), C.byref(st), C.byref(gs)))
    film.spp += spp
    return st, gs


def _batch_array(jobs):
    """jobs: dicts with scene, film, spp and optional num_path_vertices, seed, use_mis, first_sample."""
    arr = (BatchJob * len(jobs))()
    for i, j in enumerate(jobs):
        film = j.get("film")  # autotune renders into scratch films and needs none
        first = j.get("first_sample")
        if first is None:
            first = film.spp if film is not None else 0
        arr[i] = BatchJob(j["scene"].h, film.h if film is not None else None, j["spp"], j.get("num_path_vertices", 10),
                          j.get("seed", 0), int(bool(j.get("use_mis", True))), first)
    return arr


def render_batch(jobs, max_paths_in_flight: int = 0, profile: bool = False, count_traversal: bool = False,
                 use_graph: bool = False, fetch_chunk: int = 0, refill_lanes: int = 0,
                 tuning: RenderParams | None = None) -> RenderStats:
    """wr_render_batch: every job's film gets what render_nee would add, all jobs in one path pool."""
    arr = _batch_array(jobs)
    p = render_params(max_paths_in_flight=max_paths_in_flight, profile=profile, count_traversal=count_traversal,
                      use_graph=use_graph, fetch_chunk=fetch_chunk, refill_lanes=refill_lanes, tuning=tuning)
    st = RenderStats()
    _check(lib().wr_render_batch(arr, len(jobs), C.byref(p), C.byref(st)))
    for j in jobs:
        j["film"].spp += j["spp"]
    return st


def autotune(jobs) -> tuple[RenderParams, float]:
    """wr_autotune: the fastest pool size / traversal chunk / refill threshold / graph replay for a batch."""
    arr = _batch_array(jobs)
    base, best, rate = render_params(), RenderParams(), C.c_double()
    _check(lib().wr_autotune(arr, len(jobs), C.byref(base), C.byref(best), C.byref(rate)))
    return best, rate.value
