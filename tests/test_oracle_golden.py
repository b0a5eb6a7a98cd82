"""The oracle against golden vectors produced by the reference's own classes
(tests/golden/ref_vectors.json, written by tests/golden/make_golden.py)."""
import json
import os

import numpy as np
import pytest

import oracle
from wurstrenderer_amd import scenes

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.json")))
# Material parameters reach the oracle as FP32 (as they reach the GPU), so values
# proportional to a reflectance or to sigma carry FP32 rounding of the inputs.
RTOL_PARAM = 2e-7 * 8
ATOL = 1e-12


def mat_of(rec):
    m = rec["mat"]
    return oracle.material(m["type"], m["twosided"], m["rho_d"], m["rho_s"], m["sigma"], m["exponent"])


def test_onb_matches_reference():
    for rec in GOLD["onb"]:
        x, z = oracle.onb(rec["n"])
        np.testing.assert_allclose(x, rec["x"], rtol=0, atol=1e-12)
        np.testing.assert_allclose(z, rec["z"], rtol=0, atol=1e-12)


def test_mappings_match_reference():
    for rec in GOLD["mapping"]:
        np.testing.assert_allclose(oracle.cosine_hemisphere(rec["u"]), rec["cosine_hemisphere"], atol=1e-12)
        np.testing.assert_allclose(oracle.barycentric_from_square(rec["u"]), rec["barycentric"], atol=1e-12)
        np.testing.assert_allclose(oracle.disk_from_square(rec["u"]), rec["disk"], atol=1e-12)


def test_cdf_table_matches_reference():
    for rec in GOLD["cdf"]:
        i, pdf, rem = oracle.cdf_sample(rec["w"], rec["u"])
        assert i == rec["index"]
        assert pdf == pytest.approx(rec["pdf"], abs=1e-15)
        assert rem == pytest.approx(rec["remapped"], abs=1e-12)


@pytest.mark.parametrize("kind", ["bsdf_eval"])
def test_bsdf_eval_matches_reference(kind):
    n_nonzero = 0
    for rec in GOLD[kind]:
        f, pdf = oracle.bsdf_eval(mat_of(rec), rec["wi"], rec["wo"])
        np.testing.assert_allclose(f, rec["f"], rtol=RTOL_PARAM, atol=ATOL, err_msg=str(rec))
        assert pdf == pytest.approx(rec["pdf"], rel=1e-10, abs=ATOL), rec
        n_nonzero += any(v != 0 for v in rec["f"])
    assert n_nonzero > 100  # the vectors exercise non-trivial lobes


def test_bsdf_sample_matches_reference():
    n_nonzero = 0
    for rec in GOLD["bsdf_sample"]:
        wo, w, pdf = oracle.bsdf_sample(mat_of(rec), rec["wi"], rec["u"])
        np.testing.assert_allclose(wo, rec["wo"], rtol=0, atol=1e-12, err_msg=str(rec))
        np.testing.assert_allclose(w, rec["w"], rtol=RTOL_PARAM, atol=ATOL, err_msg=str(rec))
        assert pdf == pytest.approx(rec["pdf"], rel=1e-10, abs=ATOL), rec
        n_nonzero += any(v != 0 for v in rec["w"])
    assert n_nonzero > 100


def _light_scene():
    b = scenes.SceneBuilder()
    for m in GOLD["light_meshes"]:
        b.mesh(np.array(m["verts"]).reshape(-1, 3), np.array(m["tris"]).reshape(-1, 3), light_L=m["L"])
    return b.build(dict(pos=(0, 0, 5), lookat=(0, 0, 0), fovy=40.0, resolution=(8, 8)), "lights")


def test_light_sampling_matches_reference():
    sc = _light_scene()
    for m, li in zip(GOLD["light_meshes"], sc.lights):
        tris = sc.tris[sc.tri_geom == li["geom"]]
        np.testing.assert_array_equal(sc.verts[tris].astype(np.float64),
                                      np.array(m["verts"]).reshape(-1, 3)[np.array(m["tris"]).reshape(-1, 3)])
    osc = oracle.OracleScene(sc)
    for rec in GOLD["light_sample"]:
        li, pos, nrm, con, pdfA = osc.light_sample(rec["u"])
        assert li == rec["light"]
        np.testing.assert_allclose(pos, rec["pos"], rtol=0, atol=1e-12)
        np.testing.assert_allclose(nrm, rec["normal"], rtol=0, atol=1e-12)
        np.testing.assert_allclose(con, rec["contrib"], rtol=1e-12)
        assert pdfA == pytest.approx(rec["pdfA"], rel=1e-12)


def test_camera_matches_reference():
    for rec in GOLD["camera"]:
        c = rec["cam"]
        b = scenes.SceneBuilder()
        b.mesh(*scenes.quad((0, 50, 0), (1, 0, 0), (0, 0, 1)), light_L=(1, 1, 1))
        sc = b.build(dict(pos=c["pos"], lookat=c["lookat"], up=c["up"], fovy=c["fovy"],
                          resolution=(c["width"], c["height"]), aperture=c["aperture"], focal=c["focal"]), "cam")
        osc = oracle.OracleScene(sc)
        o, d, r, pdfW = osc.camera_ray(rec["pixel"][0], rec["pixel"][1], rec["ulens"], rec["upix"])
        np.testing.assert_allclose(o, rec["org"], rtol=0, atol=1e-12)
        np.testing.assert_allclose(d, rec["dir"], rtol=0, atol=1e-12)
        np.testing.assert_allclose(r, rec["raster"], rtol=0, atol=1e-9)
        assert pdfW == pytest.approx(rec["pdfW"], rel=1e-12)
        assert pdfW == pytest.approx(rec["pdfW_eval"], rel=1e-9)
        we1, r2 = osc.camera_we1(rec["org"], rec["w"])
        assert we1 == pytest.approx(rec["we1"], rel=1e-10, abs=1e-12)
        if rec["we1"] != 0:
            np.testing.assert_allclose(r2, rec["we1_raster"], rtol=0, atol=1e-8)


def test_reference_harness_when_available():
    """In a container with the reference checkout, rebuild the harness and check
    that it still reproduces the committed fixture exactly."""
    if not os.path.isdir("/root/reference/src"):
        pytest.skip("reference checkout not present")
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run(["make", "-s", "-C", os.path.join(root, "oracle"), "ref"], check=True)
    out = json.loads(subprocess.run([oracle.REF_VECTORS_BIN], check=True, capture_output=True, text=True).stdout)
    for k, v in out.items():
        assert GOLD[k] == v, k
