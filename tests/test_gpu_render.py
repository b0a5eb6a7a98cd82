"""GPU NEE+MIS renders against the FP64 oracle.

Both sides draw the same counter-based random numbers per (pixel, sample,
dimension), so at equal seeds the two images follow the same paths and agree
pixel by pixel up to FP32-vs-FP64 arithmetic; a path diverges only where FP32
rounding flips a discrete decision (a hit near an edge, a CDF boundary).
Bars, written per test: relMSE(gpu, oracle) at the same seed at least 30x
below the relMSE between two oracle images of different seeds (the estimator
noise), image means within 1e-3 relative, and extension-ray counts within
0.5%.  Convergence at independent seeds is checked against the noise level."""
import numpy as np
import pytest

import oracle
import wurstrenderer_amd as wr
from wurstrenderer_amd import scenes

pytestmark = pytest.mark.gpu


def _gpu_image(dev, sc, spp, **kw):
    gs = wr.GpuScene(dev, sc)
    film = wr.Film(gs)
    st = wr.render_nee(gs, film, spp, **kw)
    return film.image(), st, film


def _matched(dev, sc, spp, N, seed, use_mis=True, noise_factor=30.0, mean_rtol=1e-3):
    print(f"matched render: {sc.name} {sc.resolution} spp={spp} N={N} mis={use_mis}", flush=True)
    img, st, _ = _gpu_image(dev, sc, spp, num_path_vertices=N, seed=seed, use_mis=use_mis)
    osc = oracle.OracleScene(sc)
    ref, ost = osc.image(spp, num_path_vertices=N, seed=seed, use_mis=use_mis)
    other, _ = osc.image(spp, num_path_vertices=N, seed=seed + 1000, use_mis=use_mis)
    assert np.isfinite(img).all()
    noise = oracle.rel_mse(other, ref)
    diff = oracle.rel_mse(img, ref)
    print(f"  relMSE(gpu, oracle)={diff:.3e}  noise={noise:.3e}  rays gpu={st.extend_rays} oracle={ost.extend_rays}")
    assert diff * noise_factor < noise, (diff, noise)
    np.testing.assert_allclose(img.mean((0, 1)), ref.mean((0, 1)), rtol=mean_rtol)
    assert abs(st.extend_rays / ost.extend_rays - 1) < 5e-3
    assert st.shadow_rays <= ost.shadow_rays  # the GPU skips zero-contribution shadow rays
    assert st.paths == sc.resolution[0] * sc.resolution[1] * spp
    return img, ref, st


def test_cornell_small_matches_oracle(gpu_device):
    _matched(gpu_device, scenes.cornell(1, resolution=(64, 64)), 16, 10, seed=3)


def test_cornell_without_mis_matches_oracle(gpu_device):
    _matched(gpu_device, scenes.cornell(1, resolution=(64, 64)), 16, 10, seed=4, use_mis=False)


def test_cornell_full_config_matches_oracle(gpu_device):
    """The CPU-reference workload: 256x256, 16 spp, max depth 8 (N = 10)."""
    _matched(gpu_device, scenes.cornell(1), 16, 10, seed=1)


def _all_materials_scene():
    sc = scenes.cornell(7, resolution=(64, 64))
    mats = sc.materials.copy()
    # walls: Oren-Nayar; tall box: Phong; short box: Mixed (Lambert + Phong); back: one-sided Lambert
    mats[0]["type"], mats[0]["sigma"] = scenes.OREN_NAYAR, 0.5
    extra = np.zeros(3, scenes.MATERIAL_DTYPE)
    extra[0] = (scenes.PHONG, 1, (0, 0, 0), (0.7, 0.7, 0.7), 0.0, 40.0)
    extra[1] = (scenes.MIXED, 1, (0.3, 0.4, 0.5), (0.4, 0.4, 0.4), 0.0, 300.0)
    extra[2] = (scenes.LAMBERT, 0, (0.6, 0.6, 0.6), (0, 0, 0), 0.0, 0.0)
    n0 = len(mats)
    sc.materials = np.concatenate([mats, extra])
    gm = sc.geom_material.copy()
    gm[5], gm[6], gm[2] = n0, n0 + 1, n0 + 2
    sc.geom_material = gm
    return sc


def test_all_material_types_match_oracle(gpu_device):
    _matched(gpu_device, _all_materials_scene(), 16, 10, seed=5)


def test_thin_lens_camera_matches_oracle(gpu_device):
    sc = scenes.cornell(1, resolution=(48, 48)).with_camera(aperture_radius=0.05, focal_distance=3.5)
    _matched(gpu_device, sc, 16, 8, seed=6)


def test_icospheres_reduced_matches_oracle(gpu_device):
    sc = scenes.icospheres(2, n_spheres=12, level=3, resolution=(96, 54))
    _matched(gpu_device, sc, 8, 18, seed=7)


def test_full_size_scene_reduced_resolution_matches_oracle(gpu_device):
    """The 1.31M-triangle workload at 192x108, 2 spp, max depth 16 (N = 18)."""
    sc = scenes.icospheres(2, resolution=(192, 108))
    _matched(gpu_device, sc, 2, 18, seed=8)


def test_independent_seeds_converge_to_oracle(gpu_device):
    sc = scenes.cornell(1, resolution=(32, 32))
    img, _, _ = _gpu_image(gpu_device, sc, 1024, num_path_vertices=10, seed=11)
    osc = oracle.OracleScene(sc)
    ref, _ = osc.image(1024, num_path_vertices=10, seed=12)
    ref2, _ = osc.image(1024, num_path_vertices=10, seed=13)
    noise = oracle.rel_mse(ref2, ref)
    assert oracle.rel_mse(img, ref) < 1.5 * noise
    np.testing.assert_allclose(img.mean((0, 1)), ref.mean((0, 1)), rtol=0.01)


def test_progressive_calls_and_batching_are_exact(gpu_device):
    """Two calls of 8 spp equal one call of 16 spp, and splitting the paths in
    flight into several batches changes nothing but the FP64 summation order."""
    sc = scenes.cornell(1, resolution=(48, 40))
    gs = wr.GpuScene(gpu_device, sc)
    f1, f2, f3 = wr.Film(gs), wr.Film(gs), wr.Film(gs)
    wr.render_nee(gs, f1, 16, num_path_vertices=10, seed=2)
    wr.render_nee(gs, f2, 8, num_path_vertices=10, seed=2)
    wr.render_nee(gs, f2, 8, num_path_vertices=10, seed=2)
    wr.render_nee(gs, f3, 16, num_path_vertices=10, seed=2, max_paths_in_flight=48 * 40 * 3)
    b1, s1 = f1.read()
    for f in (f2, f3):
        b, s = f.read()
        np.testing.assert_allclose(b, b1, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(s, s1, rtol=1e-9, atol=1e-9)


def test_profile_and_counters(gpu_device):
    sc = scenes.cornell(1, resolution=(64, 64))
    _, st, _ = _gpu_image(gpu_device, sc, 4, num_path_vertices=10, seed=1, profile=True, count_traversal=True)
    assert st.extend_ms > 0 and st.shade_ms > 0
    # one traversal launch per pool iteration (at least N-1 of them); it traces
    # the shadow rays of the previous iteration too, so none is timed apart
    assert st.extend_launches >= 9 and st.shadow_launches == 0 and st.shadow_ms == 0
    assert st.extend_nodes >= st.extend_rays and st.extend_tris > 0 and st.shadow_nodes > 0


def test_degenerate_requests(gpu_device):
    sc = scenes.cornell(1, resolution=(16, 16))
    gs = wr.GpuScene(gpu_device, sc)
    film = wr.Film(gs)
    st = wr.render_nee(gs, film, 0, num_path_vertices=10)
    assert st.paths == 0
    st = wr.render_nee(gs, film, 4, num_path_vertices=1)  # pathnextevent.h:163: nothing to do
    assert st.paths == 0
    b, s = film.read()
    assert not b.any() and not s.any()
    with pytest.raises(wr.WrError):
        wr.render_nee(gs, film, -1)
    # an oversized work chunk is rejected before any launch (it could wrap the
    # work counter), for single renders and batches alike
    with pytest.raises(wr.WrError, match="fetch_chunk"):
        wr.render_nee(gs, film, 4, num_path_vertices=10, fetch_chunk=1 << 20)
    with pytest.raises(wr.WrError, match="fetch_chunk"):
        wr.render_batch([dict(scene=gs, film=film, spp=4, num_path_vertices=10, seed=1)], fetch_chunk=4097)
    b, s = film.read()
    assert not b.any() and not s.any()


def test_closing_a_scene_closes_its_films_and_guides(gpu_device):
    """The library needs films and guides destroyed before their scene; the
    binding closes them first, so dropping them afterwards is harmless."""
    sc = scenes.cornell(1, resolution=(16, 16))
    gs = wr.GpuScene(gpu_device, sc)
    film = wr.Film(gs)
    g = wr.Guide(gs)
    wr.render_nee(gs, film, 2, num_path_vertices=6, seed=1)
    gs.close()
    assert film.h is None and g.h is None
    del film, g
    other = wr.GpuScene(gpu_device, sc)
    f2 = wr.Film(other)
    wr.render_nee(other, f2, 2, num_path_vertices=6, seed=1)
    assert np.isfinite(f2.image()).all()
