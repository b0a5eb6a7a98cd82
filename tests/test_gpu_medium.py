"""GPU homogeneous medium (free flight, HG phase, transmittance-weighted
shadow rays) against the FP64 oracle and closed forms.

Bars, written per test: matched seeds relMSE(gpu, oracle) at least 10x below
the estimator noise (an FP32 free-flight decision near a surface can send a
path another way, so the bar is looser than the vacuum renders' 30x) and
image means within 5e-3; closed forms within the estimator's own variance;
guided sampling in a medium unbiased (means within 4 sigma)."""
import math

import numpy as np
import pytest

import oracle
import wurstrenderer_amd as wr
from wurstrenderer_amd import scenes

pytestmark = pytest.mark.gpu


def _gpu(dev, sc, spp, N, seed, use_mis=True):
    gs = wr.GpuScene(dev, sc)
    film = wr.Film(gs)
    st = wr.render_nee(gs, film, spp, num_path_vertices=N, seed=seed, use_mis=use_mis)
    return film.image(), st


@pytest.mark.parametrize("use_mis", [True, False])
def test_medium_matches_oracle(gpu_device, use_mis):
    sc = scenes.cornell(1, resolution=(48, 48)).with_medium(sigma_a=(0.05, 0.1, 0.15), sigma_s=(0.4, 0.3, 0.2), g=0.7)
    print(f"medium matched render: cornell 48x48 mis={use_mis}", flush=True)
    img, st = _gpu(gpu_device, sc, 32, 12, seed=3, use_mis=use_mis)
    osc = oracle.OracleScene(sc)
    ref, ost = osc.image(32, num_path_vertices=12, seed=3, use_mis=use_mis)
    other, _ = osc.image(32, num_path_vertices=12, seed=1003, use_mis=use_mis)
    noise, diff = oracle.rel_mse(other, ref), oracle.rel_mse(img, ref)
    print(f"  relMSE(gpu, oracle)={diff:.3e} noise={noise:.3e} rays gpu={st.extend_rays} oracle={ost.extend_rays}")
    assert diff * 10 < noise
    np.testing.assert_allclose(img.mean((0, 1)), ref.mean((0, 1)), rtol=5e-3)
    assert abs(st.extend_rays / ost.extend_rays - 1) < 5e-3


def test_absorbing_medium_closed_form(gpu_device):
    sigma = 0.6
    sc = scenes.emitting_box().with_medium(sigma_a=sigma, sigma_s=0.0, g=0.0)
    print("absorbing medium in an emitting box", flush=True)
    a, _ = _gpu(gpu_device, sc, 1024, 4, seed=1)
    b, _ = _gpu(gpu_device, sc, 1024, 4, seed=2)
    rays = scenes.camera_rays(sc, spp=64, seed=3)
    d = scenes.box_exit_distance(rays[:, :3].astype(np.float64), rays[:, 4:7].astype(np.float64))
    W, H = sc.resolution
    expect = np.exp(-sigma * d).reshape(H, W, 64).mean(axis=2)
    var = ((a[..., 0] - b[..., 0]) ** 2).mean() / 2
    err = ((a[..., 0] - expect) ** 2).mean()
    print(f"  squared error {err:.3e} vs estimator variance {var:.3e}")
    assert err < 1.3 * var + 1e-6


def test_scattering_medium_equilibrium(gpu_device):
    sc = scenes.emitting_box(resolution=(32, 32)).with_medium(sigma_a=0.0, sigma_s=0.5, g=0.7)
    print("scattering medium in an emitting box (equilibrium L = 1)", flush=True)
    img, st = _gpu(gpu_device, sc, 1024, 64, seed=4)
    noise = img[..., 0].std() / math.sqrt(img[..., 0].size)
    print(f"  mean {img.mean():.5f} noise {noise:.2e}")
    assert abs(img.mean() - 1.0) < 5 * noise + 1e-3
    assert st.shadow_rays > 0


def test_guided_sampling_in_medium_is_unbiased(gpu_device):
    sc = scenes.cornell(1, resolution=(40, 40)).with_medium(sigma_a=0.05, sigma_s=0.4, g=0.7)
    gs = wr.GpuScene(gpu_device, sc)
    print("guided (random network) vs plain in a medium", flush=True)
    g = wr.Guide(gs, seed=3)
    imgs = []
    for seed, guided in ((11, False), (12, False), (13, True), (14, True)):
        film = wr.Film(gs)
        if guided:
            wr.render_guided(gs, film, g, 1024, num_path_vertices=12, seed=seed, guide_fraction=0.5, spp_per_pass=1024,
                             train_steps=0, warmup_passes=0)
        else:
            wr.render_nee(gs, film, 1024, num_path_vertices=12, seed=seed)
        imgs.append(film.image())
    p1, p2, g1, g2 = imgs
    npx = p1.shape[0] * p1.shape[1]
    sig = np.sqrt(((p1 - p2) ** 2).mean((0, 1)) / 2 / npx + ((g1 - g2) ** 2).mean((0, 1)) / 2 / npx)
    print(f"  plain {p1.mean((0, 1))} guided {g1.mean((0, 1))} sigma {sig}")
    assert np.all(np.abs(g1.mean((0, 1)) - p1.mean((0, 1))) < 4 * sig + 1e-6)
