"""GPU BVH build (PLOC, collapsed to the 4-wide layout) against the host SAH build.

Bars, written per test:
  - invariants: every triangle in exactly one leaf, leaf records equal the
    input, child boxes enclosed by their parent entry, each node referenced
    once, stack bound within the traversal stack (the checks of the host
    builder's own self-test, run on the downloaded GPU tree);
  - determinism: two builds of one scene are byte-identical (hash);
  - quality: SAH cost within 1.3x of the host binned-SAH build;
  - hits: closest-hit t bit-identical to the host-built tree except near ties
    (fraction <= 1e-3, each within 1e-5 t: a hit within an ulp of a box face
    that one tree culls behind an equally close hit), and (geom, prim)
    identical wherever t is, except exact-t ties (shared edges: both triangles
    return the same FP64-fallback t; fraction <= 5e-3), which each tree
    resolves by its own visiting order;
  - render: the same seed renders the same image with either tree (relMSE
    between them 1e-3 of the relMSE between two seeds)."""
import numpy as np
import pytest

import oracle
import wurstrenderer_amd as wr
from wurstrenderer_amd import scenes

pytestmark = pytest.mark.gpu


def identical_triangles(n):
    b = scenes.SceneBuilder()
    m = b.material(scenes.LAMBERT, rho_d=(0.5, 0.5, 0.5))
    v = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64), (n, 1))
    b.mesh(v, np.arange(3 * n).reshape(-1, 3), material=m)
    b.mesh(*scenes.quad((0, 3, 0), (0.5, 0, 0), (0, 0, 0.5)), light_L=(1, 1, 1))
    return b.build(dict(pos=(0, 0, 5), lookat=(0, 0, 0), fovy=45.0, resolution=(16, 16)), name=f"identical_{n}")


SMALL = [("soup1", lambda: scenes.triangle_soup(1, seed=1)),
         ("soup2", lambda: scenes.triangle_soup(2, seed=2)),
         ("soup3", lambda: scenes.triangle_soup(3, seed=3)),
         ("soup7", lambda: scenes.triangle_soup(7, seed=4)),
         ("soup3000", lambda: scenes.triangle_soup(3000, seed=5)),
         ("identical1000", lambda: identical_triangles(1000)),
         ("cornell", lambda: scenes.cornell(1, resolution=(32, 32))),
         ("icospheres", lambda: scenes.icospheres(2, n_spheres=12, level=3, resolution=(32, 18)))]


@pytest.mark.parametrize("name,make", SMALL, ids=[s[0] for s in SMALL])
def test_gpu_bvh_invariants_and_determinism(gpu_device, name, make):
    sc = make()
    print(f"gpu bvh: {name} tris={len(sc.tris)}", flush=True)
    a = wr.GpuScene(gpu_device, sc)
    sa = a.bvh_check(sc)
    b = wr.GpuScene(gpu_device, sc)
    sb = b.bvh_check(sc)
    print(f"  nodes {sa['n_nodes']} depth {sa['depth']} stack {sa['max_stack']} iters {sa['iterations']} "
          f"sah {sa['sah_cost']:.2f} {sa['build_ms']:.1f} ms", flush=True)
    assert sa["builder"] == wr.BVH_GPU and sa["hash"] == sb["hash"]
    assert 1 <= sa["max_stack"] <= 64 and sa["n_nodes"] >= 1
    if len(sc.tris) > 1:
        assert sa["iterations"] >= 1


def _pair(gpu_device, sc):
    g = wr.GpuScene(gpu_device, sc)
    h = wr.GpuScene(gpu_device, sc, builder="host")
    return g, h, g.bvh_check(sc), h.bvh_check(sc)


@pytest.mark.parametrize("n_tris", [200_000, 1_000_000])
def test_gpu_bvh_quality_and_hits_match_host_build(gpu_device, n_tris):
    sc = scenes.cornell_clutter(7, n_tris, resolution=(160, 120))
    print(f"gpu vs host bvh: clutter {len(sc.tris)} tris", flush=True)
    g, h, sg, sh = _pair(gpu_device, sc)
    print(f"  gpu: {sg['build_ms']:.1f} ms, {sg['n_nodes']} nodes, depth {sg['depth']}, stack {sg['max_stack']}, "
          f"{sg['iterations']} iterations, SAH {sg['sah_cost']:.2f}")
    print(f"  host: {sh['build_ms']:.1f} ms, {sh['n_nodes']} nodes, depth {sh['depth']}, stack {sh['max_stack']}, "
          f"SAH {sh['sah_cost']:.2f}")
    assert sg["sah_cost"] <= 1.3 * sh["sah_cost"]
    rays = np.concatenate([scenes.camera_rays(sc, spp=2, seed=1), scenes.random_rays(sc, 40000, seed=2, inside=True)])
    gg, gp, gt = g.trace_closest(rays)
    hg, hp, ht = h.trace_closest(rays)
    same_t = gt.view(np.uint32) == ht.view(np.uint32)
    print(f"  rays {len(rays)}: t differs on {(~same_t).sum()}, prim differs at equal t on "
          f"{(same_t & ((gg != hg) | (gp != hp))).sum()}")
    diff = ~same_t
    assert diff.mean() <= 1e-3
    assert np.all(np.isfinite(gt[diff]) & np.isfinite(ht[diff]))
    assert np.all(np.abs(gt[diff] - ht[diff]) <= 1e-5 * ht[diff]), np.abs(gt[diff] - ht[diff]) / ht[diff]
    assert (same_t & ((gg != hg) | (gp != hp))).mean() <= 5e-3
    occ_g, occ_h = g.trace_any(rays), h.trace_any(rays)
    assert (occ_g == occ_h).mean() >= 1 - 1e-4


def test_render_with_gpu_and_host_bvh_agree(gpu_device):
    sc = scenes.cornell_clutter(8, 100_000, resolution=(64, 48))
    g, h, _, _ = _pair(gpu_device, sc)
    print("render gpu vs host bvh: clutter 100k 64x48", flush=True)
    imgs = []
    for gs, seed in ((g, 5), (h, 5), (g, 6)):
        film = wr.Film(gs)
        wr.render_nee(gs, film, 64, num_path_vertices=10, seed=seed)
        imgs.append(film.image())
    same, noise = oracle.rel_mse(imgs[0], imgs[1]), oracle.rel_mse(imgs[0], imgs[2])
    print(f"  relMSE gpu/host tree {same:.3e}, two seeds {noise:.3e}")
    assert same <= 1e-3 * noise


def test_builder_argument_errors(gpu_device):
    sc = scenes.cornell(1, resolution=(8, 8))
    with pytest.raises(ValueError):
        wr.GpuScene(gpu_device, sc, builder="embree")
    other = scenes.cornell_clutter(1, 2000, resolution=(8, 8))
    gs = wr.GpuScene(gpu_device, sc)
    with pytest.raises(wr.WrError):
        gs.bvh_check(other)
