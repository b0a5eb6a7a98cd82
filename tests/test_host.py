"""The C++ host (wurstrenderer_amd/wurst_amd): scene JSON + OBJ/MTL loading with
the reference's material mapping, and its error behaviour.  No GPU: --check
parses and validates, --dump writes the flattened scene."""
import json
import os
import subprocess

import numpy as np
import pytest

from wurstrenderer_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "wurstrenderer_amd", "wurst_amd")
NEE = {"integrator": "nextevent_path_tracer", "output": "out.pfm", "num_spp": 4}


def run(args, cwd=None):
    return subprocess.run([EXE] + args, capture_output=True, text=True, cwd=cwd, timeout=120)


def write(tmp_path, sc, renders=(NEE,), extra=None):
    scenes.write_obj(sc, str(tmp_path), "mesh")
    j = scenes.scene_json(sc, "mesh.obj", list(renders))
    if extra:
        j.update(extra)
    p = tmp_path / "scene.json"
    p.write_text(json.dumps(j))
    return p


def dump(tmp_path, sc):
    p = write(tmp_path, sc)
    d = tmp_path / "dump"
    d.mkdir()
    r = run(["--check", "--dump", str(d), "--simple2parse", str(p)])
    assert r.returncode == 0, r.stderr
    rd = lambda n, t: np.fromfile(d / n, dtype=t)
    return dict(verts=rd("verts.f32", np.float32).reshape(-1, 3), tris=rd("tris.u32", np.uint32).reshape(-1, 3),
                tri_geom=rd("tri_geom.u32", np.uint32), geom_material=rd("geom_material.i32", np.int32),
                geom_light=rd("geom_light.i32", np.int32), materials=rd("materials.bin", scenes.MATERIAL_DTYPE),
                lights=rd("lights.bin", scenes.LIGHT_DTYPE), camera=rd("camera.bin", scenes.CAMERA_DTYPE)[0])


def _mixed_scene():
    sc = scenes.icospheres(3, n_spheres=6, level=2, resolution=(40, 30))
    mats = sc.materials.copy()
    mats[0] = (scenes.MIXED, 1, (0.3, 0.4, 0.5), (0.25, 0.25, 0.25), 0.0, 60.0)
    sc.materials = mats
    return sc


@pytest.mark.parametrize("maker", ["cornell", "icospheres", "mixed"])
def test_obj_roundtrip_matches_generated_scene(tmp_path, maker):
    sc = {"cornell": lambda: scenes.cornell(1, resolution=(64, 48)),
          "icospheres": lambda: scenes.icospheres(2, n_spheres=8, level=2, resolution=(40, 30)),
          "mixed": _mixed_scene}[maker]()
    d = dump(tmp_path, sc)
    np.testing.assert_array_equal(d["tri_geom"], sc.tri_geom)
    np.testing.assert_array_equal(d["verts"][d["tris"]], sc.verts[sc.tris])  # same triangles, same order
    np.testing.assert_array_equal(d["geom_light"] >= 0, sc.geom_light >= 0)
    for g in range(len(sc.geom_material)):
        if sc.geom_light[g] >= 0:
            np.testing.assert_array_equal(d["lights"][d["geom_light"][g]]["L"], sc.lights[sc.geom_light[g]]["L"])
            assert d["lights"][d["geom_light"][g]]["geom"] == g
        else:
            a, b = d["materials"][d["geom_material"][g]], sc.materials[sc.geom_material[g]]
            assert a["type"] == b["type"] and a["twosided"] == 1
            if b["type"] != scenes.PHONG:
                np.testing.assert_array_equal(a["rho_d"], b["rho_d"])
            if b["type"] != scenes.LAMBERT:
                np.testing.assert_array_equal(a["rho_s"], b["rho_s"])
                assert a["exponent"] == b["exponent"]  # Ns / 4 (assimpio.h:310-315)
    assert d["camera"].tobytes() == sc.camera.tobytes()


def test_mitsuba_srgb_converts_reflectances(tmp_path):
    sc = scenes.cornell(1, resolution=(16, 16))
    p = write(tmp_path, sc)
    j = json.loads(p.read_text())
    j["meshes"][0]["mitsuba_srgb"] = True
    p.write_text(json.dumps(j))
    d = tmp_path / "dump"
    d.mkdir()
    assert run(["--check", "--dump", str(d), "--simple2parse", str(p)]).returncode == 0
    mats = np.fromfile(d / "materials.bin", dtype=scenes.MATERIAL_DTYPE)
    srgb = lambda v: v / 12.92 if v <= 0.04045 else ((v + 0.055) / 1.055) ** 2.4  # spectrum.h:13-17
    got = mats[1]["rho_d"]  # the ceiling (gray 0.75)
    np.testing.assert_allclose(got, [srgb(np.float64(np.float32(0.75)))] * 3, rtol=1e-6)


@pytest.mark.parametrize("mutate,msg", [
    (lambda j: j["render"][0].update(integrator="nonsense"), "unknown integrator"),
    (lambda j: j["render"][0].update(integrator="photon_map"), "not part of this build"),
    (lambda j: j["camera"].update(radius=0), "focal distance"),
    (lambda j: j["meshes"][0].update(path="missing.obj"), "cannot open file"),
    (lambda j: j.update(envmap="sky.hdr"), "envmap"),
    (lambda j: j["camera"].update(fovy=200), "fov"),
    (lambda j: j["camera"].update(resolution=[4]), "can't parse resolution"),
    (lambda j: j.update(medium={"type": "fog"}), "unknown medium"),
    (lambda j: j.update(medium={"type": "homogeneous", "sigma_a": [0.1, 0.1], "sigma_s": 0.4}), "can't parse sigma_a"),
    (lambda j: j.update(medium={"type": "homogeneous", "sigma_a": -1, "sigma_s": 0.4}), "must be >= 0"),
    (lambda j: j.update(medium={"type": "homogeneous", "sigma_a": 0.1, "sigma_s": 0.4, "g": 1.0}), "(-1, 1)"),
    (lambda j: j["render"][0].update(guiding={"guide_fraction": 1.5}), "guide_fraction"),
])
def test_errors_are_reported(tmp_path, mutate, msg):
    p = write(tmp_path, scenes.cornell(1, resolution=(16, 16)))
    j = json.loads(p.read_text())
    mutate(j)
    p.write_text(json.dumps(j))
    r = run(["--check", "--simple2parse", str(p)])
    assert r.returncode == 1 and msg in r.stderr, r.stderr


def test_malformed_json_and_emitter_with_exponent(tmp_path):
    r = run(["--check", "--jsimple2parse", '{"camera": [1, 2,'])
    assert r.returncode == 1 and "json" in r.stderr
    sc = scenes.cornell(1, resolution=(16, 16))
    p = write(tmp_path, sc)
    mtl = (tmp_path / "mesh.mtl").read_text().replace("Ke 17 17 17", "Ke 17 17 17\nNs 10")
    (tmp_path / "mesh.mtl").write_text(mtl)
    r = run(["--check", "--simple2parse", str(p)])
    assert r.returncode == 1 and "cone-emission" in r.stderr


def test_no_lights_is_an_error(tmp_path):
    sc = scenes.cornell(1, resolution=(16, 16))
    p = write(tmp_path, sc)
    mtl = (tmp_path / "mesh.mtl").read_text().replace("Ke 17 17 17", "Ke 0 0 0\nKd 0.5 0.5 0.5")
    (tmp_path / "mesh.mtl").write_text(mtl)
    r = run(["--check", "--simple2parse", str(p)])
    assert r.returncode == 1 and "no lights" in r.stderr


def test_default_txt_is_created_when_missing(tmp_path):
    r = run([], cwd=str(tmp_path))
    assert r.returncode == 0 and (tmp_path / "default.txt").exists()


def test_medium_and_guiding_keys_are_accepted(tmp_path):
    sc = scenes.cornell(1, resolution=(16, 16)).with_medium(sigma_a=(0.05, 0.05, 0.05), sigma_s=0.4, g=0.7)
    p = write(tmp_path, sc, renders=(dict(NEE, guiding={"guide_fraction": 0.5, "train_steps": 2, "train_spp": 4,
                                                        "overlap_training": True, "avg_decay": 0.99}),))
    j = json.loads(p.read_text())
    assert j["medium"] == {"type": "homogeneous", "sigma_a": [0.05] * 3, "sigma_s": [0.4] * 3, "g": 0.7}
    r = run(["--check", "--simple2parse", str(p)])
    assert r.returncode == 0, r.stderr
    # a single number broadcasts (Vec3FromJson, json.h:9-15); "none" is no medium
    for med in ({"type": "homogeneous", "sigma_a": 0.1, "sigma_s": 0.2}, {"type": "none"}):
        j["medium"] = med
        p.write_text(json.dumps(j))
        assert run(["--check", "--simple2parse", str(p)]).returncode == 0


def test_other_integrators_of_this_build_parse(tmp_path):
    """primitive_path_tracer, light_splatting, visualize and
    bidirectional_path_tracer are dispatched by name as the reference's parser
    does (simple2parser.h:176-199); the rest stay rejected."""
    sc = scenes.cornell(1, resolution=(16, 16))
    p = write(tmp_path, sc, renders=({"integrator": "primitive_path_tracer", "num_spp": 2, "num_path_vertices": 5},
                                     {"integrator": "light_splatting", "num_spp": 2, "num_path_vertices": 5},
                                     {"integrator": "visualize", "num_spp": 2},
                                     {"integrator": "bidirectional_path_tracer", "num_spp": 2, "num_path_vertices": 5}))
    r = run(["--check", "--simple2parse", str(p)])
    assert r.returncode == 0, r.stderr
    for name in ("photon_map", "diffuse_prt", "glossy_prt2002"):
        p = write(tmp_path, sc, renders=({"integrator": name},))
        r = run(["--check", "--simple2parse", str(p)])
        assert r.returncode != 0 and "not part of this build" in r.stderr


def test_envmap_pfm_is_loaded_and_checked(tmp_path):
    """The scene JSON's "envmap" (a .pfm, as FimageIo::LoadPfm reads it) lights a
    scene without area lights; other formats and a non-2:1 image are errors."""
    sc = scenes.icospheres(2, n_spheres=4, level=1, resolution=(16, 8))
    sc.lights = sc.lights[:0]
    sc.geom_light = np.full_like(sc.geom_light, -1)
    sc.geom_material = np.where(sc.geom_material < 0, 0, sc.geom_material).astype(np.int32)
    r = run(["--check", "--simple2parse", str(write(tmp_path, sc))])
    assert r.returncode == 1 and "no lights" in r.stderr
    sky = sc.with_envmap(np.random.default_rng(1).uniform(0.1, 2.0, (8, 16, 3)))
    r = run(["--check", "--simple2parse", str(write(tmp_path, sky))])
    assert r.returncode == 0, r.stderr
    bad = sc.with_envmap(np.ones((8, 8, 3)))
    r = run(["--check", "--simple2parse", str(write(tmp_path, bad))])
    assert r.returncode == 1 and "wrong size" in r.stderr


def test_geometry_ids_follow_the_reference_attach_order(tmp_path):
    """Per OBJ the reference attaches the surfaces first and then the emitters,
    each in file order, OBJ after OBJ (assimpio.h:123-150 splits them,
    simple2parser.h:38-61 attaches mGeometries before mLightGeometries); the
    light table follows the emitters in the same order.  The OBJs here list an
    emitter first."""
    def quad_lines(k, y):
        b = 4 * k
        return [f"v -1 {y} -1", f"v 1 {y} -1", f"v 1 {y} 1", f"v -1 {y} 1", f"f {b + 1} {b + 2} {b + 3} {b + 4}"]
    mtl = ["newmtl red", "Kd 0.5 0.1 0.1", "Ks 0 0 0", "Ns 0", "Ke 0 0 0",
           "newmtl gray", "Kd 0.4 0.4 0.4", "Ks 0 0 0", "Ns 0", "Ke 0 0 0",
           "newmtl lampA", "Kd 0 0 0", "Ks 0 0 0", "Ns 0", "Ke 1 2 3",
           "newmtl lampB", "Kd 0 0 0", "Ks 0 0 0", "Ns 0", "Ke 4 5 6",
           "newmtl lampC", "Kd 0 0 0", "Ks 0 0 0", "Ns 0", "Ke 7 8 9"]
    (tmp_path / "m.mtl").write_text("\n".join(mtl) + "\n")
    a = ["mtllib m.mtl"]
    for k, (name, y) in enumerate([("lampA", 3.0), ("red", 0.0), ("lampB", 4.0), ("gray", 1.0)]):
        ql = quad_lines(k, y)
        a += [f"o part{k}", f"usemtl {name}"] + ql[:4] + [ql[4]]
    (tmp_path / "a.obj").write_text("\n".join(a) + "\n")
    b = ["mtllib m.mtl"]
    for k, (name, y) in enumerate([("lampC", 5.0), ("gray", 2.0)]):
        ql = quad_lines(k, y)
        b += [f"o part{k}", f"usemtl {name}"] + ql[:4] + [ql[4]]
    (tmp_path / "b.obj").write_text("\n".join(b) + "\n")
    j = {"camera": {"pos": [0, 0, 6], "lookat": [0, 0, 0], "up": [0, 1, 0], "resolution": [8, 8], "fovy": 40.0},
         "meshes": [{"path": "a.obj"}, {"path": "b.obj"}], "render": [NEE]}
    p = tmp_path / "scene.json"
    p.write_text(json.dumps(j))
    d = tmp_path / "dump"
    d.mkdir()
    r = run(["--check", "--dump", str(d), "--simple2parse", str(p)])
    assert r.returncode == 0, r.stderr
    gm = np.fromfile(d / "geom_material.i32", np.int32)
    gl = np.fromfile(d / "geom_light.i32", np.int32)
    lights = np.fromfile(d / "lights.bin", scenes.LIGHT_DTYPE)
    mats = np.fromfile(d / "materials.bin", scenes.MATERIAL_DTYPE)
    verts = np.fromfile(d / "verts.f32", np.float32).reshape(-1, 3)
    tris = np.fromfile(d / "tris.u32", np.uint32).reshape(-1, 3)
    tg = np.fromfile(d / "tri_geom.u32", np.uint32)
    # a.obj: red, gray, lampA, lampB; then b.obj: gray, lampC
    np.testing.assert_array_equal(gl, [-1, -1, 0, 1, -1, 2])
    assert (gm[[0, 1, 4]] >= 0).all() and (gm[[2, 3, 5]] == -1).all()
    np.testing.assert_allclose(mats[gm[0]]["rho_d"], [0.5, 0.1, 0.1])
    np.testing.assert_allclose(mats[gm[1]]["rho_d"], [0.4, 0.4, 0.4])
    np.testing.assert_array_equal(lights["geom"], [2, 3, 5])
    np.testing.assert_array_equal(lights["L"], [[1, 2, 3], [4, 5, 6], [7, 8, 9]])
    heights = [float(verts[tris[tg == g]][..., 1].mean()) for g in range(6)]
    np.testing.assert_allclose(heights, [0.0, 1.0, 3.0, 4.0, 2.0, 5.0])
