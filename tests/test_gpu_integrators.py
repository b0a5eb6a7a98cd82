"""GPU primitive_path_tracer, light_splatting and visualize against the FP64 oracle.

Both sides draw the same counter-based random numbers per (pixel, sample,
dimension), so at equal seeds the images follow the same paths and agree pixel
by pixel up to FP32-vs-FP64 arithmetic.  Bars, written per test:
  - primitive: relMSE(gpu, oracle) at the same seed at least 30x (10x in a
    medium) below the relMSE between two oracle seeds, image means within 3e-3
    (5e-3 in a medium; without NEE one path that an FP32 rounding diverts onto
    the emitter moves its pixel by L T / spp, and the seed-to-seed noise of the
    cornell mean is ~1.7%), extension-ray counts within 0.5%, no shadow rays and
    an empty splat film; inside a box of emitting walls the image is exactly L;
  - light_splatting: relMSE(gpu, oracle) at the same seed at least 30x below
    the seed noise, means within 3e-3, extension- and shadow-ray counts within
    0.5%, an empty base film; GPU light tracing and GPU NEE+MIS agree (means
    1%); a medium changes nothing (the reference's light sampler is not
    volumetric);
  - visualize: per pixel |gpu - oracle| <= 1e-5 on at least 99% of the pixels
    (the rest are pixels where an FP32 rounding moves a sample's hit across a
    triangle edge), exact 1 on a seen emitter and 0 on misses."""
import numpy as np
import pytest

import oracle
import wurstrenderer_amd as wr
from wurstrenderer_amd import scenes

from test_gpu_render import _all_materials_scene

pytestmark = pytest.mark.gpu


def _primitive(dev, sc, spp, N, seed, noise_factor=30.0, mean_rtol=3e-3):
    print(f"primitive matched render: {sc.name} {sc.resolution} spp={spp} N={N}", flush=True)
    gs = wr.GpuScene(dev, sc)
    film = wr.Film(gs)
    st = wr.render_primitive(gs, film, spp, num_path_vertices=N, seed=seed)
    img = film.image()
    base, splat = film.read()
    osc = oracle.OracleScene(sc)
    rb, _, ost = osc.render(spp, num_path_vertices=N, seed=seed, integrator="primitive_path_tracer")
    ob, _, _ = osc.render(spp, num_path_vertices=N, seed=seed + 1000, integrator="primitive_path_tracer")
    ref, other = rb / spp, ob / spp
    noise, diff = oracle.rel_mse(other, ref), oracle.rel_mse(img, ref)
    print(f"  relMSE(gpu, oracle)={diff:.3e}  noise={noise:.3e}  rays gpu={st.extend_rays} oracle={ost.extend_rays}")
    assert np.isfinite(img).all()
    assert diff * noise_factor < noise, (diff, noise)
    np.testing.assert_allclose(img.mean((0, 1)), ref.mean((0, 1)), rtol=mean_rtol)
    assert abs(st.extend_rays / ost.extend_rays - 1) < 5e-3
    assert st.shadow_rays == 0 and not splat.any()
    return img, ref


def test_primitive_cornell_matches_oracle(gpu_device):
    _primitive(gpu_device, scenes.cornell(1, resolution=(64, 64)), 64, 10, seed=3)


def test_primitive_all_materials_match_oracle(gpu_device):
    _primitive(gpu_device, _all_materials_scene(), 64, 10, seed=5)


def test_primitive_in_medium_matches_oracle(gpu_device):
    sc = scenes.cornell(1, resolution=(48, 48)).with_medium(sigma_a=(0.05, 0.1, 0.15), sigma_s=(0.4, 0.3, 0.2), g=0.7)
    _primitive(gpu_device, sc, 64, 12, seed=7, noise_factor=10.0, mean_rtol=5e-3)


def test_primitive_emitting_box_is_exact(gpu_device):
    gs = wr.GpuScene(gpu_device, scenes.emitting_box(resolution=(16, 16), L=2.5))
    film = wr.Film(gs)
    wr.render_primitive(gs, film, 8, num_path_vertices=10, seed=1)
    assert np.all(film.image() == 2.5)


def test_primitive_and_nee_agree_on_the_gpu(gpu_device):
    sc = scenes.cornell(1, resolution=(32, 32))
    gs = wr.GpuScene(gpu_device, sc)
    fa, fb = wr.Film(gs), wr.Film(gs)
    wr.render_nee(gs, fa, 1024, num_path_vertices=10, seed=1)
    wr.render_primitive(gs, fb, 2048, num_path_vertices=10, seed=2)
    np.testing.assert_allclose(fb.image().mean((0, 1)), fa.image().mean((0, 1)), rtol=0.01)


def _visualize(dev, sc, spp, seed):
    print(f"visualize: {sc.name} {sc.resolution} spp={spp}", flush=True)
    gs = wr.GpuScene(dev, sc)
    film = wr.Film(gs)
    st = wr.render_visualize(gs, film, spp, seed=seed)
    base, _, _ = oracle.OracleScene(sc).render(spp, seed=seed, integrator="visualize")
    assert st.paths == sc.resolution[0] * sc.resolution[1] * spp
    return film.image(), base / spp


@pytest.mark.parametrize("make", [lambda: scenes.cornell(1, resolution=(64, 64)),
                                  lambda: scenes.icospheres(2, n_spheres=12, level=3, resolution=(96, 54))],
                         ids=["cornell", "icospheres"])
def test_visualize_matches_oracle(gpu_device, make):
    img, ref = _visualize(gpu_device, make(), 16, seed=2)
    err = np.abs(img - ref).max(axis=2)
    print(f"  pixels off by > 1e-5: {(err > 1e-5).mean():.4f}, max {err.max():.3e}")
    assert (err <= 1e-5).mean() >= 0.99
    assert np.abs(img - ref).mean() < 1e-3


def test_visualize_emitter_exact(gpu_device):
    b = scenes.SceneBuilder()
    b.mesh(*scenes.quad((0, 0, 0), (1, 0, 0), (0, 1, 0)), light_L=(3.0, 2.0, 1.0))
    sc = b.build(dict(pos=(0, 0, 3), lookat=(0, 0, 0), fovy=60.0, resolution=(16, 16)), "emitter")
    img, ref = _visualize(gpu_device, sc, 64, seed=4)
    assert np.all(img[5:11, 5:11] == 1.0) and np.all(img[0, 0] == 0.0)
    np.testing.assert_array_equal(img, ref)


def test_integrator_argument_errors(gpu_device):
    sc = scenes.cornell(1, resolution=(16, 16))
    ga, gb = wr.GpuScene(gpu_device, sc), wr.GpuScene(gpu_device, sc)
    fb = wr.Film(gb)
    with pytest.raises(wr.WrError):
        wr.render_visualize(ga, fb, 4)
    with pytest.raises(wr.WrError):
        wr.render_primitive(ga, fb, 4)
    fa = wr.Film(ga)
    with pytest.raises(wr.WrError):
        wr.render_visualize(ga, fa, -1)
    assert wr.render_primitive(ga, fa, 4, num_path_vertices=1).paths == 0
    assert not fa.read()[0].any()


# ----------------------------------------------------------------- light_splatting
def _light(dev, sc, spp, N, seed, noise_factor, mean_rtol=3e-3):
    print(f"light tracer matched render: {sc.name} {sc.resolution} spp={spp} N={N}", flush=True)
    W, H = sc.resolution
    gs = wr.GpuScene(dev, sc)
    film = wr.Film(gs)
    st = wr.render_light(gs, film, spp, num_path_vertices=N, seed=seed)
    base, splat = film.read()
    img = film.image()
    osc = oracle.OracleScene(sc)
    _, rs, ost = osc.render(spp, num_path_vertices=N, seed=seed, integrator="light_splatting")
    _, os_, _ = osc.render(spp, num_path_vertices=N, seed=seed + 1000, integrator="light_splatting")
    ref, other = rs / (spp * W * H), os_ / (spp * W * H)
    noise, diff = oracle.rel_mse(other, ref), oracle.rel_mse(img, ref)
    print(f"  relMSE(gpu, oracle)={diff:.3e}  noise={noise:.3e}  rays gpu={st.extend_rays}/{st.shadow_rays} "
          f"oracle={ost.extend_rays}/{ost.shadow_rays}")
    assert not base.any() and np.isfinite(img).all()
    assert diff * noise_factor < noise, (diff, noise)
    np.testing.assert_allclose(img.mean((0, 1)), ref.mean((0, 1)), rtol=mean_rtol)
    assert abs(st.extend_rays / ost.extend_rays - 1) < 5e-3
    assert abs(st.shadow_rays / ost.shadow_rays - 1) < 5e-3
    return img, ref


def test_light_tracer_cornell_matches_oracle(gpu_device):
    _light(gpu_device, scenes.cornell(1, resolution=(64, 64)), 32, 10, seed=3, noise_factor=30.0)


def test_light_tracer_all_materials_match_oracle(gpu_device):
    _light(gpu_device, _all_materials_scene(), 32, 10, seed=5, noise_factor=30.0)


def test_light_tracer_thin_lens_matches_oracle(gpu_device):
    sc = scenes.cornell(1, resolution=(48, 48)).with_camera(aperture_radius=0.05, focal_distance=3.5)
    _light(gpu_device, sc, 32, 8, seed=6, noise_factor=30.0)


def test_light_tracer_and_nee_agree_on_the_gpu(gpu_device):
    sc = scenes.cornell(1, resolution=(32, 32))
    gs = wr.GpuScene(gpu_device, sc)
    fa, fb = wr.Film(gs), wr.Film(gs)
    wr.render_nee(gs, fa, 1024, num_path_vertices=10, seed=1)
    wr.render_light(gs, fb, 1024, num_path_vertices=10, seed=2)
    np.testing.assert_allclose(fb.image().mean((0, 1)), fa.image().mean((0, 1)), rtol=0.01)
    assert oracle.rel_mse(fa.image(), fb.image()) < 5e-3


def test_light_tracer_ignores_the_medium(gpu_device):
    sc = scenes.cornell(1, resolution=(32, 32))
    ga = wr.GpuScene(gpu_device, sc)
    gb = wr.GpuScene(gpu_device, sc.with_medium(sigma_a=(0.3, 0.3, 0.3), sigma_s=(0.3, 0.3, 0.3), g=0.5))
    fa, fb = wr.Film(ga), wr.Film(gb)
    wr.render_light(ga, fa, 16, num_path_vertices=10, seed=4)
    wr.render_light(gb, fb, 16, num_path_vertices=10, seed=4)
    np.testing.assert_allclose(fb.read()[1], fa.read()[1], rtol=1e-9, atol=1e-12)  # FP64 atomics: order only
