"""Pins the oracle's bidirectional path tracer (BidirPathRenderer,
renderer/bidirpath.h:44-421).

  - the pdfs its balance-heuristic weights use: for sampled camera samples,
    every strategy's connected path is re-derived here from its vertices alone
    (positions, normals, materials, lights, camera).  The light-side area pdf of
    vertex i is the light choice times 1/area (i = 1), the cosine emission pdf
    (i = 2) or the BSDF pdf at vertex i - 1 given vertex i - 2 (BSDF values
    from the golden-pinned `oracle.bsdf_eval`); the camera-side pdf is the
    camera's film pdf (i = n - 1) or the BSDF pdf at vertex i + 1 given i + 2;
    each times |cos| / d^2 at vertex i.  Both arrays must match the oracle's
    (rtol 1e-9) and the weight must equal p_s / sum_k p_k in product form
    (p_k = prod_{i <= k} lp_i prod_{k < i < n} cp_i), so the weights of one path
    over its strategies sum to 1.  Dropping any of the four pdf re-evaluations
    at the connection, a transposed area conversion or the light choice in the
    emitter's pdf fails this;
  - the estimator: BDPT and NEE+MIS (pinned separately) agree in the image mean
    within 5 standard errors (seed-to-seed noise) on Lambert, all-material,
    glossy and multi-light scenes at several depths; at N = 2 only emission and
    its light->camera connection remain;
  - unsupported combinations are errors, N <= 1 renders nothing."""
import numpy as np
import pytest

import oracle
from wurstrenderer_amd import scenes

from test_gpu_render import _all_materials_scene


def _glossy_box():
    sc = scenes.cornell(3, resolution=(24, 24))
    mats = sc.materials.copy()
    for i in range(len(mats)):
        mats[i]["type"], mats[i]["rho_d"], mats[i]["rho_s"], mats[i]["exponent"] = scenes.PHONG, 0, 0.8, 30.0
    sc.materials = mats
    return sc


def _multi_light():
    return scenes.icospheres(2, n_spheres=8, level=2, resolution=(24, 16))


class _TruePdfs:
    """Area pdfs of a connected path from its vertices (the definitions, not the oracle's walk)."""

    def __init__(self, sc):
        self.sc = sc
        self.mats = [oracle.material(m["type"], m["twosided"], m["rho_d"], m["rho_s"], m["sigma"], m["exponent"])
                     for m in sc.materials]
        w = np.linalg.norm(sc.lights["L"].astype(np.float64), axis=1)
        self.sel = w / w.sum()
        v = sc.verts.astype(np.float64)
        t = sc.tris
        area = np.linalg.norm(np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]]), axis=1) / 2
        self.light_area = np.array([area[sc.tri_geom == g].sum() for g in sc.lights["geom"]])
        cam = sc.camera
        W, H = sc.resolution
        z = cam["lookat"].astype(np.float64) - cam["pos"].astype(np.float64)
        self.cam_z = z / np.linalg.norm(z)
        d = H * 0.5 / np.tan(np.radians(float(cam["fovy_deg"])) * 0.5)
        self.film_norm = d * d / (W * H)  # film pdf of a direction: d^2 / (W H cos^3)

    @staticmethod
    def _conv(a, b, nb):
        u = b - a
        l2 = u @ u
        return abs(nb @ u) / np.sqrt(l2) / l2

    def _bsdf_pdf(self, mat, n, at, frm, to):
        x, z = oracle.onb(n)
        loc = lambda w: np.array([w @ x, w @ n, w @ z])
        wi = frm - at
        wo = to - at
        return oracle.bsdf_eval(self.mats[mat], loc(wi / np.linalg.norm(wi)), loc(wo / np.linalg.norm(wo)))[1]

    def arrays(self, r):
        n, P, Nn, kind, ids = r["n"], r["pos"], r["nrm"], r["kind"], r["id"]
        lp, cp = np.zeros(n), np.zeros(n)
        for i in range(n - 1):  # 0-based i is vertex i + 1
            if i == 0:
                assert kind[0] == 1
                lp[0] = self.sel[ids[0]] / self.light_area[ids[0]]
            elif i == 1:
                u = P[1] - P[0]
                lp[1] = max(Nn[0] @ u / np.linalg.norm(u), 0.0) / np.pi * self._conv(P[0], P[1], Nn[1])
            else:
                assert kind[i - 1] == 2
                lp[i] = self._bsdf_pdf(ids[i - 1], Nn[i - 1], P[i - 1], P[i - 2], P[i]) * self._conv(P[i - 1], P[i], Nn[i])
            if i == n - 2:
                assert kind[n - 1] == 0
                u = P[i] - P[n - 1]
                c = self.cam_z @ (u / np.linalg.norm(u))
                cp[i] = self.film_norm / c ** 3 * self._conv(P[n - 1], P[i], Nn[i])
            else:
                assert kind[i + 1] == 2
                cp[i] = self._bsdf_pdf(ids[i + 1], Nn[i + 1], P[i + 1], P[i + 2], P[i]) * self._conv(P[i + 1], P[i], Nn[i])
        return lp, cp


def _product_weight(lp, cp, s, n):
    p = [np.prod(lp[:k]) * np.prod(cp[k:n - 1]) for k in range(n)]
    return p[s] / sum(p)


@pytest.mark.parametrize("name,make,N,samples", [("cornell", lambda: scenes.cornell(1, resolution=(16, 16)), 8, 60),
                                                 ("all_materials", _all_materials_scene, 8, 60),
                                                 ("glossy", _glossy_box, 7, 60),
                                                 ("multi_light", _multi_light, 7, 1500),  # open: most paths escape
                                                 ("thin_lens", lambda: scenes.cornell(1, resolution=(16, 16)).with_camera(
                                                     aperture_radius=0.05, focal_distance=3.5), 6, 60)])
def test_bdpt_weights_use_the_path_pdfs(name, make, N, samples):
    sc = make()
    osc = oracle.OracleScene(sc)
    truth = _TruePdfs(sc)
    W, H = sc.resolution
    rng = np.random.default_rng(5)
    checked, strategies, lights = 0, set(), set()
    for k in range(samples):
        px, py = int(rng.integers(W)), int(rng.integers(H))
        for r in osc.bdpt_records(N, 9, px, py, k):
            n, s = r["n"], r["s"]
            lp, cp = truth.arrays(r)
            np.testing.assert_allclose(r["lp"][:n - 1], lp[:n - 1], rtol=1e-9, atol=0, err_msg=f"{name} s={s} t={r['t']}")
            np.testing.assert_allclose(r["cp"][:n - 1], cp[:n - 1], rtol=1e-9, atol=0, err_msg=f"{name} s={s} t={r['t']}")
            assert r["weight"] == pytest.approx(_product_weight(lp, cp, s, n), rel=1e-9)
            assert 0.0 < r["weight"] <= 1.0
            checked += 1
            strategies.add((s, r["t"]))
            lights.add(int(r["id"][0]))
    print(f"{name}: {checked} connections, {len(strategies)} strategies, lights {sorted(lights)}")
    assert checked > 200
    if name == "multi_light":
        assert len(lights) == 2  # the light choice enters the emitter's pdf
    else:
        assert {(1, 1), (1, 2), (2, 1), (2, 2)} <= strategies
        assert any(s == 0 for s, _ in strategies) and any(s >= 3 and t >= 3 for s, t in strategies)


def _mean_sd(osc, spp, seeds, **kw):
    x1, x2 = (osc.image(spp, seed=s, **kw)[0] for s in seeds)
    n = x1[..., 0].size
    return (x1.mean((0, 1)) + x2.mean((0, 1))) / 2, np.sqrt(((x1 - x2) ** 2).mean((0, 1)) / (2 * n) / 2)


@pytest.mark.parametrize("name,make,N", [("cornell", lambda: scenes.cornell(1, resolution=(32, 32)), 2),
                                         ("cornell", lambda: scenes.cornell(1, resolution=(32, 32)), 5),
                                         ("all_materials", _all_materials_scene, 10),
                                         ("glossy", _glossy_box, 6),
                                         ("multi_light", _multi_light, 6)])
def test_bdpt_mean_matches_nee(name, make, N):
    osc = oracle.OracleScene(make())
    mb, sb = _mean_sd(osc, 128, (1, 2), num_path_vertices=N, integrator="bidirectional_path_tracer")
    mn, sn = _mean_sd(osc, 512, (3, 4), num_path_vertices=N)
    print(name, N, mb, mn, sb, sn)
    assert np.all(np.abs(mb - mn) < 5 * np.sqrt(sb ** 2 + sn ** 2) + 1e-9)


def test_bdpt_errors_and_empty():
    sc = scenes.cornell(1, resolution=(8, 8))
    osc = oracle.OracleScene(sc)
    base, splat, st = osc.render(4, num_path_vertices=1, integrator="bidirectional_path_tracer")
    assert not base.any() and not splat.any() and st.extend_rays == 0
    with pytest.raises(RuntimeError, match="64"):
        osc.render(1, num_path_vertices=65, integrator="bidirectional_path_tracer")
    # a medium is ignored, as by the reference's non-volumetric subpath sampler
    a = osc.render(4, num_path_vertices=6, seed=3, integrator="bidirectional_path_tracer")[0]
    om = oracle.OracleScene(sc.with_medium(sigma_a=(0.3,) * 3, sigma_s=(0.3,) * 3, g=0.5))
    np.testing.assert_array_equal(om.render(4, num_path_vertices=6, seed=3, integrator="bidirectional_path_tracer")[0], a)
    env = oracle.OracleScene(sc.with_envmap((1.0, 1.0, 1.0)))
    with pytest.raises(RuntimeError, match="envmap"):
        env.render(1, integrator="bidirectional_path_tracer")
    b = scenes.SceneBuilder()
    m = b.material(scenes.LAMBERT, rho_d=(0.5, 0.5, 0.5))
    b.mesh(*scenes.quad((0, 0, 0), (1, 0, 0), (0, 1, 0)), material=m)
    dark = oracle.OracleScene(b.build(dict(pos=(0, 0, 3), lookat=(0, 0, 0), fovy=60.0, resolution=(8, 8)), "dark"))
    with pytest.raises(RuntimeError, match="no lights"):
        dark.render(1, integrator="bidirectional_path_tracer")
