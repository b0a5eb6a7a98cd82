"""GPU envmap light (wr_scene_set_envmap) against the FP64 oracle and closed forms.

Both sides draw the same counter-based random numbers, so at equal seeds the
images follow the same paths up to FP32-vs-FP64 arithmetic (texel selection by
an FP32 CdfTable can pick a neighbouring texel at a boundary).  Bars, written
per test: relMSE(gpu, oracle) at the same seed at least 10x below the relMSE
between two oracle seeds and image means within 5e-3 (NEE+MIS, MIS off,
primitive); a constant sky seen directly is L for primitive paths to FP32
rounding of the bilinear weights (1e-6) and L within 1% for NEE+MIS; a Lambert plane under it is rho L within 1%; visualize
counts a miss as 1; batches render what single calls render to FP32 rounding
(1e-6: a batch with an envmap runs other kernel instances); unsupported
combinations are errors."""
import numpy as np
import pytest

import oracle
import wurstrenderer_amd as wr
from wurstrenderer_amd import scenes

from test_oracle_envmap import _open_scene, _sky_scene, _textured_env

pytestmark = pytest.mark.gpu


def _matched(dev, sc, spp, N, seed, integrator="nextevent_path_tracer", use_mis=True, noise_factor=10.0):
    print(f"envmap matched render: {sc.name} {sc.resolution} spp={spp} N={N} {integrator} mis={use_mis}", flush=True)
    gs = wr.GpuScene(dev, sc)
    film = wr.Film(gs)
    if integrator == "primitive_path_tracer":
        st = wr.render_primitive(gs, film, spp, num_path_vertices=N, seed=seed)
    else:
        st = wr.render_nee(gs, film, spp, num_path_vertices=N, seed=seed, use_mis=use_mis)
    img = film.image()
    osc = oracle.OracleScene(sc)
    W, H = sc.resolution
    b1, s1, ost = osc.render(spp, num_path_vertices=N, seed=seed, use_mis=use_mis, integrator=integrator)
    b2, s2, _ = osc.render(spp, num_path_vertices=N, seed=seed + 1000, use_mis=use_mis, integrator=integrator)
    ref, other = b1 / spp + s1 / (spp * W * H), b2 / spp + s2 / (spp * W * H)
    noise, diff = oracle.rel_mse(other, ref), oracle.rel_mse(img, ref)
    print(f"  relMSE(gpu, oracle)={diff:.3e} noise={noise:.3e} rays gpu={st.extend_rays}/{st.shadow_rays} "
          f"oracle={ost.extend_rays}/{ost.shadow_rays}")
    assert np.isfinite(img).all()
    assert diff * noise_factor < noise, (diff, noise)
    np.testing.assert_allclose(img.mean((0, 1)), ref.mean((0, 1)), rtol=5e-3)
    assert abs(st.extend_rays / ost.extend_rays - 1) < 5e-3
    return img, ref


def test_envmap_nee_matches_oracle(gpu_device):
    _matched(gpu_device, _open_scene(), 64, 6, seed=3)


def test_envmap_nee_without_mis_matches_oracle(gpu_device):
    _matched(gpu_device, _open_scene(), 64, 6, seed=4, use_mis=False)


def test_envmap_primitive_matches_oracle(gpu_device):
    _matched(gpu_device, _open_scene(), 64, 6, seed=5, integrator="primitive_path_tracer")


def test_envmap_with_area_lights_cornell_matches_oracle(gpu_device):
    # the closed box's open front lets the sky in; the ceiling light stays
    sc = scenes.cornell(1, resolution=(48, 48)).with_envmap(_textured_env(seed=7, W=16))
    _matched(gpu_device, sc, 32, 10, seed=6)


def test_constant_sky_closed_forms(gpu_device):
    L = (0.5, 1.0, 2.0)
    gs = wr.GpuScene(gpu_device, _sky_scene(L))
    f = wr.Film(gs)
    wr.render_primitive(gs, f, 8, num_path_vertices=6, seed=1)
    # every path sees the sky at weight 1: L up to FP32 rounding of the bilinear weights
    np.testing.assert_allclose(f.image(), np.broadcast_to(np.float32(L), (16, 16, 3)), rtol=1e-6, atol=0)
    f2 = wr.Film(gs)
    wr.render_nee(gs, f2, 512, num_path_vertices=6, seed=2)
    np.testing.assert_allclose(f2.image().mean((0, 1)), L, rtol=0.01)
    f3 = wr.Film(gs)
    wr.render_visualize(gs, f3, 4, seed=3)
    assert np.all(f3.image() == 1.0)
    rho = 0.6
    gp = wr.GpuScene(gpu_device, _sky_scene(L, plane_rho=rho))
    f4 = wr.Film(gp)
    wr.render_nee(gp, f4, 256, num_path_vertices=10, seed=4)
    np.testing.assert_allclose(f4.image().mean((0, 1)), np.float32(rho) * np.float32(L), rtol=0.01)


def test_envmap_batch_equals_single_calls(gpu_device):
    a = wr.GpuScene(gpu_device, _open_scene())
    b = wr.GpuScene(gpu_device, scenes.cornell(1, resolution=(32, 32)))
    fa, fb, ga, gb = wr.Film(a), wr.Film(b), wr.Film(a), wr.Film(b)
    wr.render_batch([dict(scene=a, film=fa, spp=16, num_path_vertices=6, seed=3),
                     dict(scene=b, film=fb, spp=16, num_path_vertices=10, seed=4)])
    wr.render_nee(a, ga, 16, num_path_vertices=6, seed=3)
    wr.render_nee(b, gb, 16, num_path_vertices=10, seed=4)
    # same random numbers per sample; the batch's and the single call's kernel
    # variants differ in FP32 rounding only (measured 1e-7)
    for f, g in ((fa, ga), (fb, gb)):
        np.testing.assert_allclose(f.read()[0], g.read()[0], rtol=1e-6, atol=1e-9)
        np.testing.assert_allclose(f.read()[1], g.read()[1], rtol=1e-6, atol=1e-9)


def test_envmap_errors(gpu_device):
    gs = wr.GpuScene(gpu_device, scenes.cornell(1, resolution=(8, 8)))
    with pytest.raises(wr.WrError, match="wrong size"):
        gs.set_envmap(np.ones((4, 4, 3), np.float32))
    with pytest.raises(wr.WrError, match="black"):
        gs.set_envmap(np.zeros((4, 8, 3), np.float32))
    gs.set_envmap(np.ones((4, 8, 3), np.float32))
    f = wr.Film(gs)
    with pytest.raises(wr.WrError, match="light tracer"):
        wr.render_light(gs, f, 1)
    with pytest.raises(wr.WrError, match="guided"):
        wr.render_guided(gs, f, wr.Guide(gs), 1)
    gs.set_medium((0.1,) * 3, (0.1,) * 3, 0.0)
    with pytest.raises(wr.WrError, match="medium"):
        wr.render_nee(gs, f, 1)
    gs.set_medium()
    gs.set_envmap(None)
    wr.render_nee(gs, f, 1)  # back to the area light alone
    nolight = _sky_scene((1, 1, 1))
    nolight.envmap = None
    g2 = wr.GpuScene(gpu_device, nolight)
    f2 = wr.Film(g2)
    with pytest.raises(wr.WrError, match="no lights"):
        wr.render_nee(g2, f2, 1)
    wr.render_primitive(g2, f2, 2)
    assert not f2.read()[0].any()
