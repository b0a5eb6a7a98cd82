"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/wurst_amd.h, reports scene errors as status codes (where
the reference throws or asserts), and builds BVHs whose invariants hold."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import wurstrenderer_amd as wr
from wurstrenderer_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wurst_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    names = declared_functions()
    assert len(names) >= 15
    lib = C.CDLL(wr.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(wr.EXPORTS), set(names) ^ set(wr.EXPORTS)
    assert b"gfx950" in wr.lib().wr_version()


def test_invalid_device_ordinal_is_an_error_code():
    h = C.c_void_p()
    rc = wr.lib().wr_device_create(4096, C.byref(h))
    assert rc != 0 and not h.value
    assert wr.lib().wr_last_error()


def _expect_invalid(sc, fragment):
    with pytest.raises(wr.WrError) as e:
        wr.validate(sc)
    assert e.value.code == 1 and fragment in str(e.value), str(e.value)


def test_scene_validation_errors():
    good = scenes.cornell(1, resolution=(16, 16))
    wr.validate(good)
    bad = scenes.cornell(1, resolution=(16, 16))
    bad.tris = bad.tris.copy()
    bad.tris[3, 1] = bad.verts.shape[0]
    _expect_invalid(bad, "vertex index out of range")
    nolight = scenes.cornell(1, resolution=(16, 16))
    nolight.lights = nolight.lights[:0]
    nolight.geom_light = np.full_like(nolight.geom_light, -1)
    nolight.geom_material = np.where(nolight.geom_material < 0, 0, nolight.geom_material).astype(np.int32)
    wr.validate(nolight)  # valid: an envmap may light it; rendering NEE without any light is an error
    nomat = scenes.cornell(1, resolution=(16, 16))
    nomat.geom_material = nomat.geom_material.copy()
    nomat.geom_material[0] = 99
    _expect_invalid(nomat, "material")
    focal = good.with_camera(focal_distance=0.0)  # the JSON "radius": 0 case (thinlens.h:22-23)
    _expect_invalid(focal, "focal")
    order = scenes.cornell(1, resolution=(16, 16))
    order.tri_geom = order.tri_geom[::-1].copy()
    _expect_invalid(order, "contiguous")


def test_render_params_validation():
    """Scheduling fields out of range are errors before any launch: a fetch chunk
    above 4096 could wrap the 32-bit work counter (entries handed out twice)."""
    good = wr.RenderParams(spp=4, num_path_vertices=10, seed=1, use_mis=1, fetch_chunk=4096, refill_lanes=64)
    wr.validate_params(good)
    wr.validate_params(wr.RenderParams(spp=0, num_path_vertices=10))  # zeros select the defaults
    for field, value, fragment in [("fetch_chunk", 4097, "fetch_chunk"), ("fetch_chunk", 1 << 31, "fetch_chunk"),
                                   ("refill_lanes", 65, "refill_lanes"), ("spp", -1, "spp"),
                                   ("first_sample", -3, "first_sample"),
                                   ("max_paths_in_flight", -1, "max_paths_in_flight")]:
        p = wr.RenderParams(spp=4, num_path_vertices=10)
        setattr(p, field, value)
        with pytest.raises(wr.WrError) as e:
            wr.validate_params(p)
        assert e.value.code == 1 and fragment in str(e.value), str(e.value)
    assert wr.lib().wr_render_params_validate(None) == 1


def test_zero_area_light_is_rejected_at_scene_creation_checks():
    b = scenes.SceneBuilder()
    m = b.material(scenes.LAMBERT, rho_d=(0.5, 0.5, 0.5))
    b.mesh(*scenes.quad((0, 0, 0), (1, 0, 0), (0, 0, 1)), material=m)
    b.mesh(np.zeros((3, 3)), np.array([[0, 1, 2]]), light_L=(1, 1, 1))
    sc = b.build(dict(pos=(0, 1, 3), lookat=(0, 0, 0), fovy=40.0, resolution=(8, 8)), "degenerate")
    wr.validate(sc)  # structurally valid; the area check needs the light table built at creation
    with pytest.raises(RuntimeError, match="zero area"):
        import oracle
        oracle.OracleScene(sc)


def test_non_finite_vertices_are_rejected():
    sc = scenes.cornell(1, resolution=(8, 8))
    sc.verts = sc.verts.copy()
    sc.verts[5, 1] = np.nan
    _expect_invalid(sc, "non-finite vertex")
    sc.verts[5, 1] = np.inf
    _expect_invalid(sc, "non-finite vertex")


def test_gpu_build_entry_points_exported():
    names = set(declared_functions())
    assert {"wr_scene_create_ex", "wr_scene_bvh_check"} <= names
    assert wr.BVH_GPU == 0 and wr.BVH_HOST == 1
    assert C.sizeof(wr.BvhStats) == 48 and C.sizeof(wr.BuildParams) == 8


@pytest.mark.parametrize("maker", ["cornell", "soup", "icospheres_small", "plane"])
def test_bvh_invariants(maker):
    sc = {
        "cornell": lambda: scenes.cornell(1),
        "soup": lambda: scenes.triangle_soup(5000, seed=2),
        "icospheres_small": lambda: scenes.icospheres(2, n_spheres=12, level=3),
        "plane": lambda: scenes.plane_under_light(),
    }[maker]()
    info = wr.bvh_selftest(sc)
    assert info["nodes"] >= 1
    assert info["max_stack"] <= 64
    assert info["depth"] >= 1


def test_bvh_invariants_full_size_scene():
    """The 1.31M-triangle icosphere workload builds a valid BVH within the stack bound."""
    sc = scenes.icospheres(2)
    assert sc.n_tris == 1310726
    info = wr.bvh_selftest(sc)
    assert info["max_stack"] <= 64
