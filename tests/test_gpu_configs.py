"""The benchmark configurations' own workloads on the GPU, each checked against
the oracle at a size the oracle finishes in seconds and, at full size, through
properties that do not depend on size.

  - batched workload (32 seeded Cornell boxes with 10k..2M-triangle icosphere
    clutter, 128^2..1024^2, 16..256 spp, depth 8..16): one autotuned batch and
    one hipGraph-replayed batch each give every job exactly the film of its
    single render (FP64 summation order aside: rtol 1e-9); the batch traces
    exactly the rays of the single renders; every film is finite and >= 0;
    each job's ray counts obey the per-path bounds (1..N-1 extension rays,
    at most N-1 shadow rays per path); the two smallest scenes of the workload
    match the oracle at equal seeds at 48x48, 32 spp: relMSE 10x below the
    seed noise, more than 5% of pixels equal to 1e-4 and means within 2%.  A
    path that FP32 rounding sends another way than FP64 moves its pixel by a
    whole sample; in this closed box of interpenetrating clutter, paths of up
    to 17 bounces split that way often enough that the ratio sits at 14 at any
    spp and 13% of pixels stay identical (the same with the previous traversal
    code; a Cornell box gives 50x and 45%, the open icosphere scene 48000x and
    99%), where a systematic error would leave almost no pixel equal;
  - gap-lit room (1,003,550 triangles, 1920x1080, max depth 16): guided at
    256 spp with the benchmarked schedule (first 5 bounces guided, online
    training over the first 64 spp), has the plain renderer's mean
    (per-pixel differences average to zero within 5 standard errors), a finite
    non-negative image, and the trained network's loss gradient on records
    built from the room's own camera-ray hits matches the FP64 oracle
    (relative error < 1e-2 per layer);
  - the medium workload (seed-5 batched-workload scene, 716,836 triangles, in a
    medium of sigma_s 0.4, sigma_a 0.05, HG g 0.7; 1280x720, 512 spp, max depth
    32): guided BF16 and FP8 inference each have the plain renderer's mean (5
    standard errors), finite non-negative images; plain matches the oracle on
    the same scene at 48x27, 64 spp (relMSE 10x below the seed noise, the
    medium bar)."""
import numpy as np
import pytest

import oracle
from oracle import guide as G
import wurstrenderer_amd as wr
from wurstrenderer_amd import scenes

pytestmark = pytest.mark.gpu


def _mean_agrees(img, ref, k=5.0):
    """Per-pixel differences of two unbiased estimates of one image average to
    zero: |mean| < k standard errors, per channel."""
    d = (img - ref).reshape(-1, 3)
    se = d.std(0) / np.sqrt(len(d))
    m = d.mean(0)
    return np.all(np.abs(m) < k * se + 1e-7), m, se


def _oracle_parity(dev, sc, spp, N, seed, bar):
    gs = wr.GpuScene(dev, sc)
    film = wr.Film(gs)
    wr.render_nee(gs, film, spp, num_path_vertices=N, seed=seed)
    img = film.image()
    osc = oracle.OracleScene(sc)
    ref, _ = osc.image(spp, num_path_vertices=N, seed=seed)
    other, _ = osc.image(spp, num_path_vertices=N, seed=seed + 1000)
    diff, noise = oracle.rel_mse(img, ref), oracle.rel_mse(other, ref)
    # pixels whose every path took the oracle's way agree to FP32 rounding; a
    # path that rounding sends elsewhere moves its pixel by a whole sample
    same = (np.abs(img - ref) <= 1e-4 * (np.abs(ref) + 1e-3)).all(-1).mean()
    print(f"  oracle parity {sc.name} {sc.resolution} spp={spp} N={N}: relMSE(gpu, oracle)={diff:.3e} "
          f"noise={noise:.3e} ({noise / max(diff, 1e-30):.1f}x); pixels agreeing to 1e-4: {same:.3f}", flush=True)
    assert diff * bar < noise
    assert same > 0.05  # a systematic error would leave almost no pixel equal
    np.testing.assert_allclose(img.mean((0, 1)), ref.mean((0, 1)), rtol=2e-2)
    gs.close()


def test_batched_workload_full_size(gpu_device):
    work = scenes.batch_workload(seed=3, n_scenes=32)
    print(f"batched workload: {len(work)} scenes, {sum(s.n_tris for s, _, _ in work)} triangles", flush=True)
    gsc = [wr.GpuScene(gpu_device, sc) for sc, _, _ in work]
    jobs = [dict(scene=g, spp=spp, num_path_vertices=N, seed=1000 + i) for i, (g, (sc, spp, N)) in
            enumerate(zip(gsc, work))]
    tuning, _ = wr.autotune(jobs)
    print(f"  autotuned: pool {tuning.max_paths_in_flight} chunk {tuning.fetch_chunk} refill {tuning.refill_lanes} "
          f"graph {tuning.use_graph}", flush=True)
    for j in jobs:
        j["film"] = wr.Film(j["scene"])
    sb = wr.render_batch(jobs, tuning=tuning)
    graph_jobs = [dict(j, film=wr.Film(j["scene"])) for j in jobs]
    sg = wr.render_batch(graph_jobs, use_graph=True)
    ext = shd = 0
    for i, (j, gj, (sc, spp, N)) in enumerate(zip(jobs, graph_jobs, work)):
        f = wr.Film(j["scene"])
        st = wr.render_nee(j["scene"], f, spp, num_path_vertices=N, seed=1000 + i)
        ext += st.extend_rays
        shd += st.shadow_rays
        assert st.paths <= st.extend_rays <= st.paths * (N - 1), (i, st.paths, st.extend_rays)
        assert st.shadow_rays <= st.paths * (N - 1), (i, st.paths, st.shadow_rays)
        bs, ss = f.read()
        for what, film in (("tuned", j["film"]), ("graph", gj["film"])):
            b, s = film.read()
            assert np.isfinite(b).all() and np.isfinite(s).all() and b.min() >= 0 and s.min() >= 0, (i, what)
            np.testing.assert_allclose(b, bs, rtol=1e-9, atol=1e-12, err_msg=f"job {i} {what}: base")
            np.testing.assert_allclose(s, ss, rtol=1e-9, atol=1e-9, err_msg=f"job {i} {what}: splat")
        assert bs.sum() > 0
        for film in (f, j["film"], gj["film"]):
            film.close()
        if i > 0:  # the first scene holds the batch's path pool
            j["scene"].close()
    print(f"  batch rays {sb.extend_rays}+{sb.shadow_rays}, single renders {ext}+{shd}", flush=True)
    assert (sb.extend_rays, sb.shadow_rays) == (ext, shd)
    assert (sg.extend_rays, sg.shadow_rays) == (ext, shd)
    jobs[0]["scene"].close()
    # the two smallest scenes of the workload against the oracle at a reduced size
    for i in np.argsort([sc.n_tris for sc, _, _ in work])[:2]:
        sc, spp, N = work[i]
        t = int(sc.name.split("_t")[-1])
        small = scenes.cornell_clutter(1000 + int(i), t, (48, 48))
        assert small.n_tris == sc.n_tris
        _oracle_parity(gpu_device, small, 32, N, 1000 + int(i), 10)


def _room_records(gs, sc, n, rng):
    """Training-record inputs at the room's own surfaces: hits of random camera
    rays (position, geometric normal facing the ray, direction back along the
    ray), a cosine-distributed direction about the normal and a positive weight."""
    cam = sc.camera
    pos_c = np.array(cam["pos"], np.float64)
    look = np.array(cam["lookat"], np.float64) - pos_c
    look /= np.linalg.norm(look)
    d = look + 0.6 * rng.normal(size=(4 * n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((len(d), 8), np.float32)
    rays[:, :3] = pos_c
    rays[:, 3] = 1e-4
    rays[:, 4:7] = d
    rays[:, 7] = np.inf
    g, p, t = gs.trace_closest(rays)
    ok = g != wr.WR_INVALID_ID
    d, g, p, t = d[ok][:n], g[ok][:n], p[ok][:n], t[ok][:n]
    assert len(d) == n
    first = np.searchsorted(sc.tri_geom, g)  # triangles of a geometry are contiguous
    tri = sc.tris[first + p]
    v = sc.verts.astype(np.float64)
    ng = np.cross(v[tri[:, 1]] - v[tri[:, 0]], v[tri[:, 2]] - v[tri[:, 0]])
    ng /= np.linalg.norm(ng, axis=1, keepdims=True)
    ng *= -np.sign((ng * d).sum(1))[:, None]
    pos = pos_c + d * t[:, None]
    u1, u2 = rng.random(n), rng.random(n)
    local = np.stack([np.cos(2 * np.pi * u2) * np.sqrt(1 - u1), np.sqrt(u1), np.sin(2 * np.pi * u2) * np.sqrt(1 - u1)], 1)
    a = np.where(np.abs(ng[:, :1]) > 0.9, [[0.0, 1.0, 0.0]], [[1.0, 0.0, 0.0]])
    tx = np.cross(a, ng)
    tx /= np.linalg.norm(tx, axis=1, keepdims=True)
    tz = np.cross(tx, ng)
    wi = local[:, :1] * tx + local[:, 1:2] * ng + local[:, 2:3] * tz
    f32 = lambda x: np.ascontiguousarray(x, np.float32)
    return f32(pos), f32(ng), f32(-d), f32(wi), f32(rng.random(n) + 0.1)


def test_gap_room_full_size_guided(gpu_device):
    sc = scenes.guide_room(4, resolution=(1920, 1080))
    assert sc.n_tris == 1_003_550
    print(f"gap room {sc.n_tris} triangles, 1920x1080, guided 256 spp vs plain 256 spp", flush=True)
    gs = wr.GpuScene(gpu_device, sc)
    g = wr.Guide(gs, learning_rate=1e-2, batch_size=1 << 18, record_capacity=1 << 20, seed=7)
    film = wr.Film(gs)
    # the benchmarked schedule: the first 5 bounces guided with probability 0.4,
    # online training over the first 64 spp (64 Adam steps per 8-spp pass)
    kw = dict(num_path_vertices=18, seed=51, guide_fraction=0.4, max_guided_bounce=5)
    _, gst = wr.render_guided(gs, film, g, 64, spp_per_pass=8, train_steps=64, warmup_passes=1, **kw)
    wr.render_guided(gs, film, g, 192, spp_per_pass=64, train_steps=0, warmup_passes=0, **kw)
    guided = film.image()
    pf = wr.Film(gs)
    wr.render_nee(gs, pf, 256, num_path_vertices=18, seed=52)
    plain = pf.image()
    print(f"  trained: {gst.train_steps} Adam steps on {gst.records} records, loss {gst.last_loss:.4f}", flush=True)
    assert gst.train_steps == 8 * 64 and np.isfinite(gst.last_loss)
    assert np.isfinite(guided).all() and guided.min() >= 0
    ok, m, se = _mean_agrees(guided, plain)
    print(f"  mean guided - plain {m} (standard error {se}); image mean {plain.mean((0, 1))}", flush=True)
    assert ok, (m, se)
    # the trained network's gradient on the room's own surfaces
    rng = np.random.default_rng(61)
    pos, nrm, wo, wi, w = _room_records(gs, sc, 1 << 13, rng)
    p = g.get_params()
    loss, grad = g.gradient(pos, nrm, wo, wi, w)
    v32 = sc.verts.astype(np.float32)  # the scene bounds the network's position encoding uses
    box = np.concatenate([v32.min(0), v32.max(0)]).astype(np.float64)
    rloss, rgrad = G.loss_and_grad(p.astype(np.float64), pos, nrm, wo, wi, w, box, bf=True)
    print(f"  room records: loss gpu {loss:.6f} oracle {rloss:.6f}", flush=True)
    assert abs(loss - rloss) < 1e-3 * abs(rloss)
    o = 0
    for l, (k, n) in enumerate(G.SHAPES):
        for name, sz in (("W", k * n), ("b", n)):
            a, b = grad[o:o + sz], rgrad[o:o + sz]
            e = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            assert e < 1e-2, (l, name, e)
            o += sz
    g.close()
    gs.close()


def test_medium_workload_full_size(gpu_device):
    sc = scenes.medium_clutter(5, resolution=(1280, 720))
    print(f"medium workload {sc.name} {sc.n_tris} triangles, 1280x720, 512 spp, N=34", flush=True)
    gs = wr.GpuScene(gpu_device, sc)
    pf = wr.Film(gs)
    st = wr.render_nee(gs, pf, 512, num_path_vertices=34, seed=71)
    plain = pf.image()
    assert np.isfinite(plain).all() and plain.min() >= 0 and st.extend_rays > st.paths
    for fp8 in (False, True):
        g = wr.Guide(gs, learning_rate=1e-2, batch_size=1 << 18, record_capacity=1 << 20, seed=7)
        f = wr.Film(gs)
        wr.render_guided(gs, f, g, 128, num_path_vertices=34, seed=72, guide_fraction=0.5, spp_per_pass=8,
                         train_steps=32, warmup_passes=1, fp8_inference=fp8)
        wr.render_guided(gs, f, g, 384, num_path_vertices=34, seed=72, guide_fraction=0.5, spp_per_pass=128,
                         train_steps=0, warmup_passes=0, fp8_inference=fp8)
        img = f.image()
        assert np.isfinite(img).all() and img.min() >= 0
        ok, m, se = _mean_agrees(img, plain)
        print(f"  {'fp8' if fp8 else 'bf16'}: mean guided - plain {m} (standard error {se})", flush=True)
        assert ok, (fp8, m, se)
        f.close()
        g.close()
    gs.close()
    # the same scene at a size the oracle renders in seconds
    _oracle_parity(gpu_device, scenes.medium_clutter(5, resolution=(48, 27)), 64, 34, 73, 10)
