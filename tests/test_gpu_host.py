"""The command line renders a scene JSON on the GPU through PathNeeRenderer and
writes the reference's PFM; the image equals the Python API's render of the
same scene and seed, and matches the oracle on the same random numbers."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
import wurstrenderer_amd as wr
from wurstrenderer_amd import scenes

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "wurstrenderer_amd", "wurst_amd")


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = (int(x) for x in f.readline().split())
        assert float(f.readline()) < 0  # little endian
        data = np.frombuffer(f.read(), dtype="<f4").reshape(h, w, 3)
    return data[::-1]  # rows are stored bottom-up (floatimage.h:453-457)


def test_cli_render_equals_api_and_oracle(tmp_path, gpu_device):
    sc = scenes.cornell(1, resolution=(64, 48))
    scenes.write_obj(sc, str(tmp_path), "mesh")
    j = scenes.scene_json(sc, "mesh.obj", [{"integrator": "nextevent_path_tracer", "output": "out.pfm",
                                            "num_spp": 8, "num_path_vertices": 10, "seed": 3}])
    (tmp_path / "scene.json").write_text(json.dumps(j))
    r = subprocess.run([EXE, "--simple2parse", str(tmp_path / "scene.json")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    stats = json.loads(r.stdout.strip().splitlines()[-1])
    assert stats["paths"] == 64 * 48 * 8
    img = read_pfm(tmp_path / "out.pfm")
    gs = wr.GpuScene(gpu_device, sc)
    film = wr.Film(gs)
    wr.render_nee(gs, film, 8, num_path_vertices=10, seed=3)
    api = film.image().astype(np.float32)
    np.testing.assert_allclose(img, api, rtol=1e-6, atol=1e-7)
    ref, _ = oracle.OracleScene(sc).image(8, num_path_vertices=10, seed=3)
    other, _ = oracle.OracleScene(sc).image(8, num_path_vertices=10, seed=4)
    assert oracle.rel_mse(img, ref) * 30 < oracle.rel_mse(other, ref)


def test_cli_guided_schedule_renders_the_same_mean(tmp_path, gpu_device):
    """"guiding" with train_spp: the host trains on the first train_spp samples
    per pixel (passes of spp_per_pass, train_steps Adam steps after each), then
    renders the rest in one pass with the trained network.  Every sample is
    counted once, the image is finite and non-negative, and its mean is the
    plain estimator's (the guided estimator is unbiased for any network)."""
    sc = scenes.cornell(1, resolution=(64, 48))
    scenes.write_obj(sc, str(tmp_path), "mesh")
    guiding = {"guide_fraction": 0.4, "spp_per_pass": 2, "train_steps": 16, "train_spp": 4, "max_guided_bounce": 5,
               "record_capacity": 65536, "batch_size": 8192, "seed": 11}
    j = scenes.scene_json(sc, "mesh.obj", [{"integrator": "nextevent_path_tracer", "output": "out.pfm", "num_spp": 32,
                                            "num_path_vertices": 8, "seed": 5, "guiding": guiding}])
    (tmp_path / "scene.json").write_text(json.dumps(j))
    r = subprocess.run([EXE, "--simple2parse", str(tmp_path / "scene.json")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    stats = json.loads(r.stdout.strip().splitlines()[-1])
    assert stats["paths"] == 64 * 48 * 32
    img = read_pfm(tmp_path / "out.pfm")
    assert np.all(np.isfinite(img)) and img.min() >= 0.0
    gs = wr.GpuScene(gpu_device, sc)
    film = wr.Film(gs)
    wr.render_nee(gs, film, 256, num_path_vertices=8, seed=9)
    plain = film.image()
    assert abs(img.mean() / plain.mean() - 1.0) < 0.03, (img.mean(), plain.mean())


def test_cli_primitive_and_visualize_equal_api(tmp_path, gpu_device):
    """One scene JSON with the other integrators of this build: each PFM equals
    the Python API's render of the same scene, integrator and seed."""
    sc = scenes.cornell(2, resolution=(48, 40))
    scenes.write_obj(sc, str(tmp_path), "mesh")
    j = scenes.scene_json(sc, "mesh.obj", [
        {"integrator": "primitive_path_tracer", "output": "prim.pfm", "num_spp": 8, "num_path_vertices": 10, "seed": 5},
        {"integrator": "light_splatting", "output": "light.pfm", "num_spp": 4, "num_path_vertices": 8, "seed": 7},
        {"integrator": "visualize", "output": "vis.pfm", "num_spp": 4, "seed": 6},
        {"integrator": "bidirectional_path_tracer", "output": "bdpt.pfm", "num_spp": 4, "num_path_vertices": 7,
         "seed": 9}])
    (tmp_path / "scene.json").write_text(json.dumps(j))
    r = subprocess.run([EXE, "--simple2parse", str(tmp_path / "scene.json")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [json.loads(x) for x in r.stdout.strip().splitlines()[-4:]]
    assert [x["integrator"] for x in lines] == ["primitive_path_tracer", "light_splatting", "visualize",
                                                "bidirectional_path_tracer"]
    assert lines[0]["paths"] == 48 * 40 * 8 and lines[2]["paths"] == 48 * 40 * 4
    gs = wr.GpuScene(gpu_device, sc)
    fp, fv = wr.Film(gs), wr.Film(gs)
    wr.render_primitive(gs, fp, 8, num_path_vertices=10, seed=5)
    wr.render_visualize(gs, fv, 4, seed=6)
    np.testing.assert_allclose(read_pfm(tmp_path / "prim.pfm"), fp.image().astype(np.float32), rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(read_pfm(tmp_path / "vis.pfm"), fv.image().astype(np.float32), rtol=1e-6, atol=1e-7)
    fl = wr.Film(gs)
    wr.render_light(gs, fl, 4, num_path_vertices=8, seed=7)  # FP64 splat atomics: summation order only
    np.testing.assert_allclose(read_pfm(tmp_path / "light.pfm"), fl.image().astype(np.float32), rtol=1e-5, atol=1e-6)
    fb = wr.Film(gs)
    wr.render_bdpt(gs, fb, 4, num_path_vertices=7, seed=9)  # t = 1 splats: FP64 atomics, summation order only
    np.testing.assert_allclose(read_pfm(tmp_path / "bdpt.pfm"), fb.image().astype(np.float32), rtol=1e-5, atol=1e-6)


def test_cli_envmap_scene_equals_api(tmp_path, gpu_device):
    """A scene JSON with an "envmap" .pfm and no area light renders through the
    command line as through the Python API (same image to FP32 rounding)."""
    sc = scenes.icospheres(2, n_spheres=6, level=2, resolution=(40, 24))
    sc.lights = sc.lights[:0]
    sc.geom_light = np.full_like(sc.geom_light, -1)
    sc.geom_material = np.where(sc.geom_material < 0, 0, sc.geom_material).astype(np.int32)
    sc = sc.with_envmap(np.random.default_rng(2).uniform(0.1, 2.0, (16, 32, 3)).astype(np.float32))
    scenes.write_obj(sc, str(tmp_path), "mesh")
    j = scenes.scene_json(sc, "mesh.obj", [{"integrator": "nextevent_path_tracer", "output": "out.pfm",
                                            "num_spp": 8, "num_path_vertices": 6, "seed": 3}])
    (tmp_path / "scene.json").write_text(json.dumps(j))
    r = subprocess.run([EXE, "--simple2parse", str(tmp_path / "scene.json")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    gs = wr.GpuScene(gpu_device, sc)
    film = wr.Film(gs)
    wr.render_nee(gs, film, 8, num_path_vertices=6, seed=3)
    np.testing.assert_allclose(read_pfm(tmp_path / "out.pfm"), film.image().astype(np.float32), rtol=1e-5, atol=1e-6)
