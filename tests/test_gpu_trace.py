"""GPU traversal parity: closest-hit and any-hit against the FP64 oracle.

Bar: hit (geomID, primID) bit-exact with the oracle, misses identical, t within
the FP32 error of the watertight test, excluding documented ties:
  - another triangle hits, or nearly hits (min barycentric >= -1e-5), within
    a relative 1e-5 of the closest t (coplanar / shared-edge candidates), or
  - the closest hit lies within barycentric 1e-5 of an edge;
for a miss, a near hit anywhere on the ray.  t tolerance: |dt| <= 1e-5 t +
4e-7 C, with C the largest coordinate magnitude (the FP32 rounding of
vertex - origin, a few ulp of C, dominates for short rays from surfaces)."""
import numpy as np
import pytest

import oracle
import wurstrenderer_amd as wr
from wurstrenderer_amd import scenes

pytestmark = pytest.mark.gpu
INV = 0xFFFFFFFF


def _ties(ot, sec, mar, og):
    near_second = np.isfinite(sec) & (sec <= np.where(np.isfinite(ot), ot * (1 + 1e-5), np.inf))
    edge = (og != INV) & (mar < 1e-5)
    return near_second | edge


def _check_closest(gs, osc, rays, sc, max_tie_frac=0.01):
    print(f"closest parity: {sc.name} rays={len(rays)}", flush=True)
    g, p, t = gs.trace_closest(rays)
    og, op, ot, sec, mar = osc.trace_closest(rays, brute=True)
    tie = _ties(ot, sec, mar, og)
    ok = ~tie
    bad = ok & ((g != og) | (p != op))
    assert bad.sum() == 0, (np.nonzero(bad)[0][:10], g[bad][:5], og[bad][:5])
    hit = ok & (og != INV)
    C = float(np.abs(sc.verts).max())
    tol = 1e-5 * ot[hit] + 4e-7 * C
    assert (np.abs(t[hit] - ot[hit]) <= tol).all()
    assert tie.mean() <= max_tie_frac, tie.mean()
    assert np.isinf(t[ok & (og == INV)]).all()
    return hit.mean()


def test_closest_soup(gpu_device):
    sc = scenes.triangle_soup(3000, seed=11)
    gs, osc = wr.GpuScene(gpu_device, sc), oracle.OracleScene(sc)
    rays = np.concatenate([scenes.random_rays(sc, 20000, seed=1), scenes.random_rays(sc, 10000, seed=2, inside=True)])
    frac = _check_closest(gs, osc, rays, sc)
    assert 0.3 < frac < 0.98


def test_closest_cornell_camera_rays(gpu_device):
    sc = scenes.cornell(1)
    gs, osc = wr.GpuScene(gpu_device, sc), oracle.OracleScene(sc)
    rays = scenes.camera_rays(sc, spp=1, seed=3)
    frac = _check_closest(gs, osc, rays, sc)
    assert frac > 0.99  # the box is closed except the front


def test_closest_icospheres_random(gpu_device):
    sc = scenes.icospheres(2, n_spheres=12, level=3, resolution=(64, 36))
    gs, osc = wr.GpuScene(gpu_device, sc), oracle.OracleScene(sc)
    rays = np.concatenate([scenes.random_rays(sc, 6000, seed=4), scenes.random_rays(sc, 4000, seed=5, inside=True)])
    _check_closest(gs, osc, rays, sc)


def test_closest_full_scene_camera_rays(gpu_device):
    """The 1.31M-triangle workload: all primary rays of a 480x270 frame against
    the oracle's BVH, and a brute-force subset with tie classification."""
    sc = scenes.icospheres(2, resolution=(480, 270))
    gs, osc = wr.GpuScene(gpu_device, sc), oracle.OracleScene(sc)
    rays = scenes.camera_rays(sc, spp=1, seed=6)
    print("full scene: gpu", flush=True)
    g, p, t = gs.trace_closest(rays)
    og, op, ot = osc.trace_closest(rays)
    mism = (g != og) | (p != op)
    assert mism.mean() < 1e-4, mism.sum()
    same = ~mism & (og != INV)
    assert (np.abs(t[same] - ot[same]) <= 1e-5 * ot[same] + 4e-7 * float(np.abs(sc.verts).max())).all()
    sub = rays[np.random.default_rng(0).choice(len(rays), 300, replace=False)]
    _check_closest(gs, osc, sub, sc, max_tie_frac=0.05)


def _tiny_far_soup(n, seed, offset):
    """Triangles of ~0.02 in a 2-unit cube placed far from the origin: near the
    leaves a traversal node's quantisation grid is finer than the FP32 spacing
    at its origin, so its decoded planes snap to representable floats."""
    rng = np.random.default_rng(seed)
    b = scenes.SceneBuilder()
    m = b.material(scenes.LAMBERT, rho_d=(0.5, 0.5, 0.5))
    c = rng.uniform(-1.0, 1.0, size=(n, 1, 3))
    v = c + rng.normal(scale=0.02, size=(n, 3, 3)) + np.asarray(offset)
    b.mesh(v.reshape(-1, 3), np.arange(3 * n).reshape(-1, 3), material=m)
    b.mesh(*scenes.quad(np.asarray(offset) + (0, 3, 0), (0.5, 0, 0), (0, 0, 0.5)), light_L=(1, 1, 1))
    cam = dict(pos=tuple(np.asarray(offset) + (0, 0, 5)), lookat=tuple(offset), fovy=45.0, resolution=(16, 16))
    return b.build(cam, f"tiny_far_{n}")


def test_quantised_nodes_tiny_triangles_far_from_origin(gpu_device):
    """The traversal nodes store 8-bit child planes; their decoded boxes must
    contain the exact ones wherever FP32 is coarse relative to the boxes.
    Closest and any hit against FP64 brute force, BVH containment checked."""
    for offset in ((3.0e4, -2.0e4, 1.0e4), (-7.5e5, 0.0, 2.5e5)):
        sc = _tiny_far_soup(3000, 21, offset)
        gs, osc = wr.GpuScene(gpu_device, sc), oracle.OracleScene(sc)
        gs.bvh_check(sc)
        rays = np.concatenate([scenes.random_rays(sc, 15000, seed=22), scenes.random_rays(sc, 15000, seed=23, inside=True)])
        _check_closest(gs, osc, rays, sc, max_tie_frac=0.05)
        rays[:, 7] = np.random.default_rng(3).uniform(0.01, 3.0, len(rays))
        occ = gs.trace_any(rays)
        oocc, amb = osc.trace_any(rays, brute=True)
        ok = amb == 0
        assert (occ[ok] == oocc[ok]).all(), (offset, np.nonzero(occ[ok] != oocc[ok])[0][:10])
        assert amb.mean() < 0.05


def test_any_hit_soup_and_cornell_shadow_rays(gpu_device):
    sc = scenes.triangle_soup(3000, seed=12)
    gs, osc = wr.GpuScene(gpu_device, sc), oracle.OracleScene(sc)
    rays = np.concatenate([scenes.random_rays(sc, 20000, seed=7), scenes.random_rays(sc, 10000, seed=8, inside=True)])
    rays[:, 7] = np.random.default_rng(1).uniform(0.05, 10.0, len(rays))
    occ = gs.trace_any(rays)
    oocc, amb = osc.trace_any(rays, brute=True)
    ok = amb == 0
    assert (occ[ok] == oocc[ok]).all()
    assert 0.05 < oocc[ok].mean() < 0.9
    assert amb.mean() < 0.01
    # NEE-like rays in the Cornell box: surface points to points on the light
    sc = scenes.cornell(1)
    gs, osc = wr.GpuScene(gpu_device, sc), oracle.OracleScene(sc)
    cam = scenes.camera_rays(sc, spp=1, seed=9)[::7]
    g, p, t = osc.trace_closest(cam)
    hit = g != INV
    pos = cam[hit, :3].astype(np.float64) + cam[hit, 4:7] * t[hit, None]
    rng = np.random.default_rng(2)
    lp = np.stack([rng.uniform(-0.25, 0.25, hit.sum()), np.full(hit.sum(), 0.999), rng.uniform(-0.25, 0.25, hit.sum())], 1)
    d = lp - pos
    ln = np.linalg.norm(d, axis=1)
    sh = np.zeros((hit.sum(), 8), np.float32)
    sh[:, :3] = pos
    sh[:, 3] = 1e-4
    sh[:, 4:7] = d / ln[:, None]
    sh[:, 7] = ln - 1e-4
    occ = gs.trace_any(sh)
    oocc, amb = osc.trace_any(sh, brute=True)
    ok = amb == 0
    assert (occ[ok] == oocc[ok]).all()
    assert 0.05 < oocc[ok].mean() < 0.95
    assert amb.mean() < 0.02


def test_watertight_closed_icosphere_far_from_origin(gpu_device):
    """Rays from inside a closed mesh placed far from the origin must all hit
    (the property of Embree's WatertightTest, tutorials/verify/verify.cpp:2711-2792)."""
    v, f = scenes.icosphere(4)
    b = scenes.SceneBuilder()
    m = b.material(scenes.LAMBERT, rho_d=(0.5, 0.5, 0.5))
    centre = np.array([1000.0, -500.0, 250.0])
    b.mesh(v * 2.0 + centre, f, material=m)
    b.mesh(*scenes.quad(centre + [0, 50, 0], (1, 0, 0), (0, 0, 1)), light_L=(1, 1, 1))
    sc = b.build(dict(pos=(0, 0, 0), lookat=(1, 0, 0), fovy=40.0, resolution=(8, 8)), "ico_far")
    gs = wr.GpuScene(gpu_device, sc)
    rng = np.random.default_rng(5)
    n = 200000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = centre + rng.uniform(-0.5, 0.5, size=(n, 3))
    rays[:, 4:7] = d
    rays[:, 7] = np.inf
    g, p, t = gs.trace_closest(rays)
    assert (g == 0).all(), (g != 0).sum()
    # rays through vertices and edges of the mesh: aim exactly at vertex positions
    verts = (v * 2.0 + centre).astype(np.float32)
    o = np.repeat(centre[None].astype(np.float32), len(verts), 0)
    dd = verts.astype(np.float64) - centre
    dd /= np.linalg.norm(dd, axis=1, keepdims=True)
    r2 = np.zeros((len(verts), 8), np.float32)
    r2[:, :3] = o
    r2[:, 4:7] = dd
    r2[:, 7] = np.inf
    g2, _, _ = gs.trace_closest(r2)
    assert (g2 == 0).all()


def test_watertight_shared_edges(gpu_device):
    """Rays aimed at points on the edges two triangles share, from inside a
    closed mesh far from the origin, must all hit: the two triangles' edge
    functions for that edge are exact negations of each other only when neither
    is computed with a fused multiply-add (a crack otherwise lets a ray within an
    ulp of the edge miss both)."""
    v, f = scenes.icosphere(4)
    b = scenes.SceneBuilder()
    m = b.material(scenes.LAMBERT, rho_d=(0.5, 0.5, 0.5))
    centre = np.array([1000.0, -500.0, 250.0])
    b.mesh(v * 2.0 + centre, f, material=m)
    b.mesh(*scenes.quad(centre + [0, 50, 0], (1, 0, 0), (0, 0, 1)), light_L=(1, 1, 1))
    sc = b.build(dict(pos=(0, 0, 0), lookat=(1, 0, 0), fovy=40.0, resolution=(8, 8)), "ico_far_edges")
    gs = wr.GpuScene(gpu_device, sc)
    verts = sc.verts.astype(np.float64)
    tris = sc.tris[sc.tri_geom == 0]
    e = np.sort(np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]]), axis=1)
    edges = np.unique(e, axis=0)
    assert len(edges) == len(tris) * 3 // 2  # closed: every edge shared by two triangles
    rng = np.random.default_rng(17)
    fr = np.concatenate([[0.5], rng.uniform(0.02, 0.98, 7)])
    pts = (verts[edges[:, 0]][:, None, :] * (1 - fr)[None, :, None] +
           verts[edges[:, 1]][:, None, :] * fr[None, :, None]).reshape(-1, 3)
    origins = np.concatenate([centre[None], centre + rng.uniform(-0.6, 0.6, size=(3, 3))]).astype(np.float32)
    rays = []
    for o in origins:
        d = pts - o.astype(np.float64)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        r = np.zeros((len(pts), 8), np.float32)
        r[:, :3] = o
        r[:, 4:7] = d
        r[:, 7] = np.inf
        rays.append(r)
    rays = np.concatenate(rays)
    print(f"shared-edge rays: {len(rays)} ({len(edges)} edges x {len(fr)} points x {len(origins)} origins)", flush=True)
    g, _, t = gs.trace_closest(rays)
    assert (g == 0).all(), (g != 0).sum()
    assert gs.trace_any(rays).all()


def _hip():
    """The HIP runtime the library links (already loaded in this process; found
    through the process's own mappings so no second runtime is loaded)."""
    import ctypes as C
    wr.lib()
    path = next(line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64.so" in line)
    hip = C.CDLL(path)
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def test_closest_hit_matches_closed_form_quads(gpu_device):
    """The FP32 watertight test against closed-form ray/quad hits (pins it
    directly, not only through the oracle): hit or miss and the quad and
    triangle exact away from edges, t within the bar above."""
    from closed_form import quad_scene_and_rays
    sc, rays, geom, prim, t, amb = quad_scene_and_rays()
    print(f"closed-form quads: rays={len(rays)}", flush=True)
    gs = wr.GpuScene(gpu_device, sc)
    g, p, tt = gs.trace_closest(rays)
    ok = ~amb
    np.testing.assert_array_equal(g[ok], geom[ok])
    hit = ok & (geom != INV)
    np.testing.assert_array_equal(p[hit], prim[hit])
    C = float(np.abs(sc.verts).max())
    assert (np.abs(tt[hit] - t[hit]) <= 1e-5 * t[hit] + 4e-7 * C).all()
    assert np.isinf(tt[ok & (geom == INV)]).all()
    # any hit: occluded exactly when the closed-form closest hit lies before tmax
    occ_rays = rays.copy()
    occ_rays[:, 7] = np.random.default_rng(8).uniform(0.5, 10.0, len(rays))
    occ = gs.trace_any(occ_rays)
    sure = ok & (np.abs(t - occ_rays[:, 7]) > 1e-5 * occ_rays[:, 7])
    np.testing.assert_array_equal(occ[sure].astype(bool), (t < occ_rays[:, 7])[sure])


def test_device_resident_entry_point_matches_host_one(gpu_device):
    """wr_trace_*_device on device buffers (allocated with the library's own HIP
    runtime) gives the same hits as the host-buffer entry points."""
    import ctypes as C
    hip = _hip()
    H2D, D2H = 1, 2
    sc = scenes.triangle_soup(2000, seed=13)
    gs = wr.GpuScene(gpu_device, sc)
    rays = scenes.random_rays(sc, 50000, seed=3)
    g, p, t = gs.trace_closest(rays)
    n = len(rays)
    d_rays, d_hits, d_occ = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(d_rays), rays.nbytes) == 0
    assert hip.hipMalloc(C.byref(d_hits), 8 * n) == 0
    assert hip.hipMalloc(C.byref(d_occ), n) == 0
    try:
        assert hip.hipMemcpy(d_rays, rays.ctypes.data, rays.nbytes, H2D) == 0
        ms = C.c_float()
        wr._check(wr.lib().wr_trace_closest_device(gs.h, d_rays, n, d_hits, C.byref(ms)))
        hits = np.zeros((n, 2), np.uint32)
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(hits.ctypes.data, d_hits, 8 * n, D2H) == 0
        t_dev = hits[:, 1].view(np.float32)
        np.testing.assert_array_equal(np.isinf(t_dev), np.isinf(t))
        np.testing.assert_array_equal(t_dev[np.isfinite(t)], t[np.isfinite(t)])
        assert ((hits[:, 0] == INV) == (g == INV)).all()
        assert ms.value > 0
        wr._check(wr.lib().wr_trace_any_device(gs.h, d_rays, n, d_occ, None))
        occ = np.zeros(n, np.uint8)
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(occ.ctypes.data, d_occ, n, D2H) == 0
        np.testing.assert_array_equal(occ, gs.trace_any(rays))
    finally:
        for ptr in (d_rays, d_hits, d_occ):
            hip.hipFree(ptr)
