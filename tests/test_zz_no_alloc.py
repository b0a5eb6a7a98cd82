"""wr_render_nee_noalloc (the allocation-free NEE + MIS entry point) against
wr_render_nee, through the C-ABI.

The two render the same samples with the same random numbers, so a frame whose
spp fits one pass in both gives bit-identical FP64 base films (per-pixel sums
in sample order); splat films are FP64 atomics, whose order varies, so they
agree to FP64 rounding.  Two calls in a row, the second on a smaller frame,
then a frame that the fixed region renders in two passes (FP64 sums of two
partial passes: equal to rounding), then the size limits, which are errors
before any launch."""
import ctypes as C

import numpy as np
import pytest

import wurstrenderer_amd as wr
from wurstrenderer_amd import scenes

pytestmark = pytest.mark.gpu


def _noalloc(gs, film, spp, N, seed, **kw):
    p = wr.render_params(spp, N, seed, True, film.spp, **kw)
    st = wr.RenderStats()
    wr._check(wr.lib().wr_render_nee_noalloc(gs.h, film.h, C.byref(p), C.byref(st)))
    film.spp += spp
    return st


def _pair(gpu_device, sc, spp, N, seed):
    gs = wr.GpuScene(gpu_device, sc)
    fa, fb = wr.Film(gs), wr.Film(gs)
    sa = wr.render_nee(gs, fa, spp, num_path_vertices=N, seed=seed)
    print(f"no-alloc: {sc.name} {sc.resolution} spp={spp} N={N}", flush=True)
    sb = _noalloc(gs, fb, spp, N, seed)
    (ba, pa), (bb, pb) = fa.read(), fb.read()
    return gs, (ba, pa, sa), (bb, pb, sb)


def test_noalloc_matches_render_nee_two_calls(gpu_device):
    # first call: 64 x 64 at 16 spp (65,536 samples: one pass in both entry points)
    sc1 = scenes.icospheres(2, n_spheres=8, level=3, resolution=(64, 64))
    gs1, (ba, pa, sa), (bb, pb, sb) = _pair(gpu_device, sc1, 16, 10, 3)
    np.testing.assert_array_equal(bb, ba)
    np.testing.assert_allclose(pb, pa, rtol=1e-12, atol=1e-300)
    assert (sb.extend_rays, sb.shadow_rays, sb.paths) == (sa.extend_rays, sa.shadow_rays, sa.paths)
    assert ba.any()  # (no emitter of this frame is seen from the camera: its splat film stays 0)
    # second call, smaller input: a 32 x 24 Cornell box at 4 spp
    sc2 = scenes.cornell(1, resolution=(32, 24))
    gs2, (ba, pa, sa), (bb, pb, sb) = _pair(gpu_device, sc2, 4, 10, 5)
    np.testing.assert_array_equal(bb, ba)
    np.testing.assert_allclose(pb, pa, rtol=1e-12, atol=1e-300)
    assert (sb.extend_rays, sb.shadow_rays) == (sa.extend_rays, sa.shadow_rays)
    assert ba.any() and pa.any()  # the ceiling light is seen: t = 1 splats


def test_noalloc_multi_pass_frame(gpu_device):
    """256 x 256 at 32 spp = 2^21 samples: the fixed region renders it in two
    passes of 16 spp with 2^17 paths in flight; wr_render_nee in one pass."""
    sc = scenes.cornell(1, resolution=(256, 256))
    gs, (ba, pa, sa), (bb, pb, sb) = _pair(gpu_device, sc, 32, 10, 7)
    np.testing.assert_allclose(bb, ba, rtol=1e-13, atol=0)
    np.testing.assert_allclose(pb, pa, rtol=1e-12, atol=1e-300)
    assert (sb.extend_rays, sb.shadow_rays) == (sa.extend_rays, sa.shadow_rays)


def test_noalloc_limits_are_errors(gpu_device):
    big = scenes.cornell(1, resolution=(1100, 1000))  # 1.1M pixels > 2^20 samples per pass
    gs = wr.GpuScene(gpu_device, big)
    f = wr.Film(gs)
    with pytest.raises(wr.WrError, match="exceeds the allocation-free pass capacity"):
        _noalloc(gs, f, 1, 10, 1)
    small = scenes.cornell(1, resolution=(32, 32))
    gs2 = wr.GpuScene(gpu_device, small)
    f2 = wr.Film(gs2)
    with pytest.raises(wr.WrError, match="max_paths_in_flight"):
        _noalloc(gs2, f2, 4, 10, 1, max_paths_in_flight=1 << 18)
    with pytest.raises(wr.WrError, match="profile"):
        _noalloc(gs2, f2, 4, 10, 1, profile=True)
    with pytest.raises(wr.WrError, match="use_graph"):
        _noalloc(gs2, f2, 4, 10, 1, use_graph=True)
    f2.spp = 0
    b, s = f2.read()
    assert not b.any() and not s.any()
    # the scene still renders through the ordinary entry point afterwards
    wr.render_nee(gs2, f2, 2, num_path_vertices=6, seed=1)
    assert np.isfinite(f2.image()).all() and f2.image().any()
