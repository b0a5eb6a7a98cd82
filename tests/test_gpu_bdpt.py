"""GPU bidirectional path tracer (wr_render_bdpt) against the FP64 oracle's.

Both sides draw the same counter-based random numbers per (pixel, sample,
dimension) for the camera subpath and a derived stream for the light subpath,
so at equal seeds the images follow the same paths and agree pixel by pixel up
to FP32-vs-FP64 arithmetic.  Bars, written per test:
  - relMSE(gpu, oracle) at the same seed at least 30x below the relMSE between
    two oracle seeds; image means within 3e-3; extension- and shadow-ray counts
    within 0.5% (an FP32 rounding that moves a hit across an edge changes a
    path's length); with Phong lobes of exponent up to 1000 the shadow rays
    within 3%: 2.3% of the oracle's connections there carry an FP64 BSDF value
    below 1e-30, which FP32 flushes to 0 before the shadow ray (measured on the
    oracle's per-strategy records);
  - GPU BDPT and GPU NEE+MIS agree in the mean (1%);
  - one call of 2k samples equals two progressive calls of k (base 1e-12: FP64
    sums regrouped; splat 1e-9: FP64 atomics in another order);
  - a medium changes nothing; N <= 1 renders nothing; an envmap, N > 64 and a
    scene without lights are errors."""
import numpy as np
import pytest

import oracle
import wurstrenderer_amd as wr
from wurstrenderer_amd import scenes

from test_gpu_render import _all_materials_scene

pytestmark = pytest.mark.gpu


def _bdpt(dev, sc, spp, N, seed, noise_factor=30.0, mean_rtol=3e-3, shadow_rtol=5e-3):
    print(f"bdpt matched render: {sc.name} {sc.resolution} spp={spp} N={N}", flush=True)
    W, H = sc.resolution
    gs = wr.GpuScene(dev, sc)
    film = wr.Film(gs)
    st = wr.render_bdpt(gs, film, spp, num_path_vertices=N, seed=seed)
    img = film.image()
    osc = oracle.OracleScene(sc)
    rb, rs, ost = osc.render(spp, num_path_vertices=N, seed=seed, integrator="bidirectional_path_tracer")
    ob, os_, _ = osc.render(spp, num_path_vertices=N, seed=seed + 1000, integrator="bidirectional_path_tracer")
    ref, other = rb / spp + rs / (spp * W * H), ob / spp + os_ / (spp * W * H)
    noise, diff = oracle.rel_mse(other, ref), oracle.rel_mse(img, ref)
    print(f"  relMSE(gpu, oracle)={diff:.3e}  noise={noise:.3e}  rays gpu={st.extend_rays}/{st.shadow_rays} "
          f"oracle={ost.extend_rays}/{ost.shadow_rays}", flush=True)
    assert np.isfinite(img).all() and img.min() >= 0.0
    assert diff * noise_factor < noise, (diff, noise)
    np.testing.assert_allclose(img.mean((0, 1)), ref.mean((0, 1)), rtol=mean_rtol)
    assert abs(st.extend_rays / ost.extend_rays - 1) < 5e-3
    assert abs(st.shadow_rays / ost.shadow_rays - 1) < shadow_rtol
    assert st.paths == W * H * spp
    return img, ref


def test_bdpt_cornell_matches_oracle(gpu_device):
    _bdpt(gpu_device, scenes.cornell(1, resolution=(64, 64)), 16, 8, seed=3)


def test_bdpt_all_materials_match_oracle(gpu_device):
    _bdpt(gpu_device, _all_materials_scene(), 16, 8, seed=5)


def test_bdpt_thin_lens_matches_oracle(gpu_device):
    sc = scenes.cornell(1, resolution=(48, 48)).with_camera(aperture_radius=0.05, focal_distance=3.5)
    _bdpt(gpu_device, sc, 16, 6, seed=6)


def test_bdpt_multi_light_matches_oracle(gpu_device):
    _bdpt(gpu_device, scenes.icospheres(2, n_spheres=12, level=3, resolution=(96, 54)), 16, 10, seed=7, shadow_rtol=0.03)


def test_bdpt_direct_only(gpu_device):
    _bdpt(gpu_device, scenes.cornell(1, resolution=(64, 64)), 16, 2, seed=8)


def test_bdpt_and_nee_agree_on_the_gpu(gpu_device):
    sc = scenes.cornell(1, resolution=(32, 32))
    gs = wr.GpuScene(gpu_device, sc)
    fa, fb = wr.Film(gs), wr.Film(gs)
    wr.render_nee(gs, fa, 1024, num_path_vertices=10, seed=1)
    wr.render_bdpt(gs, fb, 512, num_path_vertices=10, seed=2)
    np.testing.assert_allclose(fb.image().mean((0, 1)), fa.image().mean((0, 1)), rtol=0.01)
    assert oracle.rel_mse(fb.image(), fa.image()) < 5e-3


def test_bdpt_progressive_equals_single_call(gpu_device):
    sc = scenes.cornell(1, resolution=(32, 32))
    gs = wr.GpuScene(gpu_device, sc)
    a, b = wr.Film(gs), wr.Film(gs)
    wr.render_bdpt(gs, a, 8, num_path_vertices=8, seed=4)
    wr.render_bdpt(gs, b, 4, num_path_vertices=8, seed=4)
    wr.render_bdpt(gs, b, 4, num_path_vertices=8, seed=4)
    ba, sa = a.read()
    bb, sb = b.read()
    np.testing.assert_allclose(bb, ba, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(sb, sa, rtol=1e-9, atol=1e-12)


def test_bdpt_ignores_the_medium_and_errors(gpu_device):
    sc = scenes.cornell(1, resolution=(16, 16))
    ga = wr.GpuScene(gpu_device, sc)
    gb = wr.GpuScene(gpu_device, sc.with_medium(sigma_a=(0.3, 0.3, 0.3), sigma_s=(0.3, 0.3, 0.3), g=0.5))
    fa, fb = wr.Film(ga), wr.Film(gb)
    wr.render_bdpt(ga, fa, 8, num_path_vertices=6, seed=4)
    wr.render_bdpt(gb, fb, 8, num_path_vertices=6, seed=4)
    np.testing.assert_array_equal(fb.read()[0], fa.read()[0])
    np.testing.assert_allclose(fb.read()[1], fa.read()[1], rtol=1e-9, atol=1e-12)
    fc = wr.Film(ga)
    st = wr.render_bdpt(ga, fc, 4, num_path_vertices=1)
    assert st.paths == 0 and not fc.read()[0].any() and not fc.read()[1].any()
    with pytest.raises(wr.WrError):
        wr.render_bdpt(ga, fc, 1, num_path_vertices=65)
    ge = wr.GpuScene(gpu_device, sc.with_envmap((1.0, 1.0, 1.0)))
    with pytest.raises(wr.WrError):
        wr.render_bdpt(ge, wr.Film(ge), 1)
    b = scenes.SceneBuilder()
    m = b.material(scenes.LAMBERT, rho_d=(0.5, 0.5, 0.5))
    b.mesh(*scenes.quad((0, 0, 0), (1, 0, 0), (0, 1, 0)), material=m)
    dark = wr.GpuScene(gpu_device, b.build(dict(pos=(0, 0, 3), lookat=(0, 0, 0), fovy=60.0, resolution=(8, 8)), "dark"))
    with pytest.raises(wr.WrError):
        wr.render_bdpt(dark, wr.Film(dark), 1)
