"""Batched rendering, hipGraph replay and autotuned scheduling on the GPU.

A batch gives every job's film what wr_render_nee with the same parameters
gives it (the same random numbers per pixel and sample); only the grouping of
the FP64 sums (passes) and the order of the splat atomics differ, so films
agree to rtol 1e-9.  Graph replay and traversal scheduling parameters change
the order of work, never the result: same bar (two triangles at exactly the
same t resolve to the lower leaf-order index, whatever the order of leaf tests
or the other rays of a wave)."""
import numpy as np
import pytest

import wurstrenderer_amd as wr
from wurstrenderer_amd import scenes

pytestmark = pytest.mark.gpu


def _films_close(a, b, what):
    ba, sa = a.read()
    bb, sb = b.read()
    np.testing.assert_allclose(ba, bb, rtol=1e-9, atol=1e-12, err_msg=f"{what}: base")
    np.testing.assert_allclose(sa, sb, rtol=1e-9, atol=1e-9, err_msg=f"{what}: splat")
    assert ba.sum() > 0


def _jobs(dev):
    specs = [(scenes.cornell(1, resolution=(48, 48)), 16, 10, True, 3),
             (scenes.cornell_clutter(1001, 40_000, (64, 40)), 8, 12, False, 4),
             (scenes.cornell_clutter(1002, 12_000, (32, 32)), 24, 18, True, 5),
             (scenes.icospheres(2, n_spheres=4, level=3, resolution=(40, 24)), 12, 8, True, 6)]
    out = []
    for sc, spp, N, mis, seed in specs:
        gs = wr.GpuScene(dev, sc)
        out.append(dict(scene=gs, spp=spp, num_path_vertices=N, use_mis=mis, seed=seed))
    return out


def test_batch_equals_individual_renders(gpu_device):
    print("batch of 4 jobs vs individual renders", flush=True)
    jobs = _jobs(gpu_device)
    for j in jobs:
        j["film"] = wr.Film(j["scene"])
    st = wr.render_batch(jobs)
    assert st.paths == sum(j["spp"] * j["scene"].W * j["scene"].H for j in jobs)
    rays = 0
    for k, j in enumerate(jobs):
        f = wr.Film(j["scene"])
        s1 = wr.render_nee(j["scene"], f, j["spp"], num_path_vertices=j["num_path_vertices"], seed=j["seed"],
                           use_mis=j["use_mis"])
        rays += s1.extend_rays + s1.shadow_rays
        _films_close(j["film"], f, f"job {k}")
    assert st.extend_rays + st.shadow_rays == rays


def test_batch_with_small_pool_and_graph(gpu_device):
    print("batch: small pool, graph replay", flush=True)
    jobs = _jobs(gpu_device)
    ref = []
    for j in jobs:
        j["film"] = wr.Film(j["scene"])
        ref.append(wr.Film(j["scene"]))
    wr.render_batch(jobs, max_paths_in_flight=5000, use_graph=True, fetch_chunk=64, refill_lanes=16)
    for k, j in enumerate(jobs):
        wr.render_nee(j["scene"], ref[k], j["spp"], num_path_vertices=j["num_path_vertices"], seed=j["seed"],
                      use_mis=j["use_mis"])
        _films_close(j["film"], ref[k], f"job {k}")


@pytest.mark.parametrize("opts", [dict(use_graph=True), dict(fetch_chunk=256, refill_lanes=40),
                                  dict(use_graph=True, max_paths_in_flight=3000)])
def test_graph_and_tuning_do_not_change_the_image(gpu_device, opts):
    sc = scenes.cornell(1, resolution=(64, 64))
    gs = wr.GpuScene(gpu_device, sc)
    print(f"render_nee with {opts}", flush=True)
    a, b = wr.Film(gs), wr.Film(gs)
    sa = wr.render_nee(gs, a, 32, num_path_vertices=10, seed=9)
    sb = wr.render_nee(gs, b, 32, num_path_vertices=10, seed=9, **opts)
    assert sa.extend_rays == sb.extend_rays and sa.shadow_rays == sb.shadow_rays
    _films_close(a, b, str(opts))


def test_autotune_returns_a_working_configuration(gpu_device):
    print("autotune on a 2-job batch", flush=True)
    jobs = _jobs(gpu_device)[:2]
    best, rate = wr.autotune(jobs)
    print(f"  best: pool {best.max_paths_in_flight} chunk {best.fetch_chunk} refill {best.refill_lanes} "
          f"graph {best.use_graph}: {rate:.1f} Mrays/s")
    assert rate > 0
    assert best.max_paths_in_flight in (1 << 23, 1 << 24, 1 << 25, 1 << 26, 1 << 27)
    assert best.fetch_chunk in (96, 128, 256) and best.refill_lanes in (16, 32)
    ref = []
    for j in jobs:
        j["film"] = wr.Film(j["scene"])
        ref.append(wr.Film(j["scene"]))
    wr.render_batch(jobs, tuning=best)
    for k, j in enumerate(jobs):
        wr.render_nee(j["scene"], ref[k], j["spp"], num_path_vertices=j["num_path_vertices"], seed=j["seed"],
                      use_mis=j["use_mis"])
        _films_close(j["film"], ref[k], f"job {k}")


def test_batch_errors(gpu_device):
    a = wr.GpuScene(gpu_device, scenes.cornell(1, resolution=(16, 16)))
    b = wr.GpuScene(gpu_device, scenes.cornell(2, resolution=(16, 16)))
    with pytest.raises(wr.WrError, match="another scene"):
        wr.render_batch([dict(scene=a, film=wr.Film(b), spp=1)])
    with pytest.raises(wr.WrError, match="above 255"):
        wr.render_batch([dict(scene=a, film=wr.Film(a), spp=1, num_path_vertices=300)])
    with pytest.raises(wr.WrError, match="empty batch"):
        wr.render_batch([])
