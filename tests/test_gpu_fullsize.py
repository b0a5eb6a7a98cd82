"""The benchmark frame at its full size (1.31M-triangle scene, 1920 x 1080,
64 spp, depth 16), checked through properties that do not depend on size.

  - one 64-spp call equals two progressive 32-spp calls: the FP64 base film sums
    each pixel's samples in sample order (relative 1e-12); the splat film gets
    FP64 atomics (1e-9);
  - every pixel is finite and non-negative; no path makes more than N - 1
    extension rays, none makes more than N - 1 shadow rays (NEE per vertex plus
    the light->camera connection);
  - the image mean equals the FP64 oracle's mean over the same full frame at
    1 spp within 5 standard errors of that oracle estimate (its own two seeds
    give the noise);
  - primitive path tracing, light tracing and bidirectional path tracing of
    the same frame agree with NEE+MIS in the mean (2%); the bidirectional
    image is finite and non-negative and makes no more than N - 1 extension
    rays per subpath."""
import numpy as np
import pytest

import oracle
import wurstrenderer_amd as wr
from wurstrenderer_amd import scenes

pytestmark = pytest.mark.gpu

N = 18


@pytest.fixture(scope="module")
def frame(gpu_device):
    print("full-size frame: icospheres 1920x1080", flush=True)
    sc = scenes.icospheres(2, resolution=(1920, 1080))
    gs = wr.GpuScene(gpu_device, sc)
    yield sc, gs
    gs.close()


def test_full_frame_progressive_and_bounds(frame):
    sc, gs = frame
    a, b = wr.Film(gs), wr.Film(gs)
    st = wr.render_nee(gs, a, 64, num_path_vertices=N, seed=1)
    s1 = wr.render_nee(gs, b, 32, num_path_vertices=N, seed=1)
    s2 = wr.render_nee(gs, b, 32, num_path_vertices=N, seed=1)
    print(f"  64 spp: {st.extend_rays} extension, {st.shadow_rays} shadow rays in {st.wall_ms:.1f} ms", flush=True)
    ba, sa = a.read()
    bb, sb = b.read()
    np.testing.assert_allclose(bb, ba, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(sb, sa, rtol=1e-9, atol=1e-9)
    assert st.extend_rays == s1.extend_rays + s2.extend_rays and st.shadow_rays == s1.shadow_rays + s2.shadow_rays
    img = a.image()
    assert np.isfinite(img).all() and img.min() >= 0.0
    assert st.paths == 1920 * 1080 * 64
    assert st.extend_rays <= st.paths * (N - 1) and st.shadow_rays <= st.paths * (N - 1)
    assert st.extend_rays >= st.paths  # every path traces its primary ray


def test_full_frame_mean_matches_oracle(frame):
    sc, gs = frame
    f = wr.Film(gs)
    wr.render_nee(gs, f, 64, num_path_vertices=N, seed=2)
    gpu = f.image().mean((0, 1))
    osc = oracle.OracleScene(sc)
    o1, _ = osc.image(1, num_path_vertices=N, seed=11)
    o2, _ = osc.image(1, num_path_vertices=N, seed=12)
    ref = (o1.mean((0, 1)) + o2.mean((0, 1))) / 2
    se = np.sqrt(((o1 - o2) ** 2).mean((0, 1)) / (2 * o1[..., 0].size) / 2)
    print(f"  gpu mean {gpu}, oracle mean {ref} +- {se}", flush=True)
    assert np.all(np.abs(gpu - ref) < 5 * se + 1e-7)


def test_full_frame_other_integrators_agree(frame):
    sc, gs = frame
    fn, fp, fl = wr.Film(gs), wr.Film(gs), wr.Film(gs)
    wr.render_nee(gs, fn, 16, num_path_vertices=N, seed=3)
    wr.render_primitive(gs, fp, 64, num_path_vertices=N, seed=4)
    wr.render_light(gs, fl, 16, num_path_vertices=N, seed=5)
    m = fn.image().mean((0, 1))
    print(f"  means: nee {m}, primitive {fp.image().mean((0, 1))}, light {fl.image().mean((0, 1))}", flush=True)
    np.testing.assert_allclose(fp.image().mean((0, 1)), m, rtol=0.02)
    np.testing.assert_allclose(fl.image().mean((0, 1)), m, rtol=0.02)


def test_full_frame_bdpt_agrees(frame):
    sc, gs = frame
    fn, fb = wr.Film(gs), wr.Film(gs)
    wr.render_nee(gs, fn, 16, num_path_vertices=N, seed=3)
    st = wr.render_bdpt(gs, fb, 4, num_path_vertices=N, seed=6)
    img = fb.image()
    print(f"  bdpt 4 spp: {st.extend_rays} extension, {st.shadow_rays} shadow rays in {st.wall_ms:.1f} ms; "
          f"means nee {fn.image().mean((0, 1))}, bdpt {img.mean((0, 1))}", flush=True)
    assert np.isfinite(img).all() and img.min() >= 0.0
    assert st.paths == 1920 * 1080 * 4
    assert st.paths <= st.extend_rays <= st.paths * 2 * (N - 1)
    np.testing.assert_allclose(img.mean((0, 1)), fn.image().mean((0, 1)), rtol=0.02)
