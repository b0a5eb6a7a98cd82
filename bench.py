"""Benchmark: GPU FP32 wavefront NEE+MIS on the 1.31M-triangle icosphere scene.

Workload (one step): one 1920x1080 frame at 64 samples per pixel, max depth 16
(num_path_vertices = 18), on 64 seeded level-5 icospheres over a 200 x 200
floor with two 1 x 1 emitters of radiance 20 (Lambert / modified Phong 50/50).
The scene, BVH and film are resident in HBM before timing starts; each step
renders the next 64 samples of the same frame.

Metric: Mrays/s = (closest-hit rays + shadow rays traced) / wall second, over
exactly K timed steps bracketed by device synchronisation.  Also reported:
samples/s, per-kernel HIP-event times, the roofline of the dominant kernel
(closest-hit traversal) and the oracle's FP64 CPU path timed on a bounded
sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, published
NODE_BYTES, TRI_BYTES = 128, 48
EXT_RAY_BYTES = 32 + 4 + 8     # ray (o, d) in + queue index + hit out
SHD_RAY_BYTES = 48             # shadow entry (o+tmax, d+target, contribution)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--depth", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU time of the baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-count", action="store_true", help="skip the traversal-counting pass")
    ap.add_argument("--no-ttr", action="store_true", help="skip the guided-vs-plain time-to-relMSE measurement")
    ap.add_argument("--ttr-width", type=int, default=480)
    ap.add_argument("--ttr-height", type=int, default=270)
    ap.add_argument("--ttr-target", type=float, default=1e-3)
    ap.add_argument("--ttr-ref-spp", type=int, default=65536)
    ap.add_argument("--ttr-runs", type=int, default=3)
    ap.add_argument("--ttr-train-spp", type=int, default=128)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-scene measurement")
    ap.add_argument("--no-integrators", action="store_true",
                    help="skip primitive / light tracing / visualize / envmap on the headline frame")
    ap.add_argument("--batch-scenes", type=int, default=32)
    ap.add_argument("--no-medium", action="store_true", help="skip the medium BF16/FP8 inference measurement")
    ap.add_argument("--med-width", type=int, default=1280)
    ap.add_argument("--med-height", type=int, default=720)
    ap.add_argument("--med-spp", type=int, default=512)
    ap.add_argument("--med-ref-spp", type=int, default=2048)
    ap.add_argument("--med-train-spp", type=int, default=128)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 1 GiB device-to-device three times (known-size kernel for byte counters)")
    return ap.parse_args()


def calibrate_copy(nbytes=1 << 30, reps=3):
    """Device-to-device copies of a known size through the HIP runtime the
    library links, so rocprofv3 byte counters can be scaled (FETCH_SIZE and
    WRITE_SIZE of the copy kernel should each read nbytes)."""
    import ctypes as C
    import wurstrenderer_amd as wr
    wr.lib()
    path = next(line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64.so" in line)
    hip = C.CDLL(path)
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), nbytes) == 0 and hip.hipMalloc(C.byref(dst), nbytes) == 0
    assert hip.hipMemset(src, 1, nbytes) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(dst, src, nbytes, 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(f"# calibration: {reps} x {nbytes} B device-to-device copies", file=sys.stderr, flush=True)


def rel_mse(img, ref, eps=1e-2):
    """mean over pixels and channels of (I - R)^2 / (R^2 + eps)"""
    import numpy as np
    return float(np.mean((img - ref) ** 2 / (ref * ref + eps)))


def progressive(render, film, ref, target, first_spp, max_spp):
    """Progressive spp doubling: wall time of the render calls until the first
    checkpoint with relMSE <= target, and the crossing time interpolated between
    the last two checkpoints (relMSE ~ 1/time there)."""
    import math
    t_total, spp, step, curve = 0.0, 0, first_spp, []
    while spp < max_spp:
        t = time.perf_counter()
        render(film, step)
        t_total += time.perf_counter() - t
        spp += step
        e = rel_mse(film.image(), ref)
        curve.append([spp, round(t_total, 4), e])
        if e <= target:
            t_cross = t_total
            if len(curve) > 1:
                (_, t0, e0), (_, t1, e1) = curve[-2], curve[-1]
                if e0 > e1 > 0:
                    t_cross = math.exp(math.log(t0) + (math.log(t1) - math.log(t0)) * math.log(e0 / target) / math.log(e0 / e1))
            return t_total, t_cross, spp, curve
        step = spp
    return None, None, spp, curve


def time_to_relmse(wr, scenes, dev, a):
    """Guided vs plain GPU NEE+MIS on the gap-lit room (1,003,550 triangles, light
    L = 500 hidden behind a 0.2 x 0.2 gap), max depth 16: wall time until the
    progressively doubled image reaches relMSE <= target against a plain
    reference at ttr_ref_spp; median of ttr_runs runs, each from an untrained
    network.  Guided: 8-spp passes with 32 Adam steps of 2^18 records after each
    until ttr_train_spp, then the trained network without further training."""
    import numpy as np
    N = 18
    sc = scenes.guide_room(4, resolution=(a.ttr_width, a.ttr_height))
    gs = wr.GpuScene(dev, sc)
    t = time.perf_counter()
    half = a.ttr_ref_spp // 2
    fa, fb = wr.Film(gs), wr.Film(gs)
    wr.render_nee(gs, fa, half, num_path_vertices=N, seed=9001)
    wr.render_nee(gs, fb, half, num_path_vertices=N, seed=9002)
    ia, ib = fa.image(), fb.image()
    ref = 0.5 * (ia + ib)
    ref_s = time.perf_counter() - t
    floor = rel_mse(ia, ib) / 4.0  # relMSE of the mean of two halves against the truth, approximately
    print(f"# ttr reference: {a.ttr_ref_spp} spp in {ref_s:.1f}s, noise floor ~{floor:.2e}", file=sys.stderr, flush=True)
    max_spp = a.ttr_ref_spp

    def plain(seed):
        return lambda film, spp: wr.render_nee(gs, film, spp, num_path_vertices=N, seed=seed)

    def guided(seed):
        g = wr.Guide(gs, learning_rate=1e-2, batch_size=1 << 18, record_capacity=1 << 20, seed=seed)
        state = {"calls": 0}

        def render(film, spp):
            # online training: 8-spp passes, 32 Adam steps after each, for the first
            # TRAIN_SPP samples per pixel; then the trained network alone
            left = max(0, a.ttr_train_spp - film.spp)
            if left:
                k = min(spp, left)
                wr.render_guided(gs, film, g, k, num_path_vertices=N, seed=seed, guide_fraction=0.5,
                                 spp_per_pass=8, train_steps=32, warmup_passes=1 if state["calls"] == 0 else 0)
                spp -= k
            if spp:
                wr.render_guided(gs, film, g, spp, num_path_vertices=N, seed=seed, guide_fraction=0.5,
                                 spp_per_pass=spp, train_steps=0, warmup_passes=0)
            state["calls"] += 1
        return render

    res = {}
    for name, make, first in (("plain", plain, 16), ("guided", guided, 16)):
        times, crosses, curves = [], [], []
        for r in range(a.ttr_runs):
            film = wr.Film(gs)
            tt, tc, spp, curve = progressive(make(100 + r), film, ref, a.ttr_target, first, max_spp)
            times.append(tt)
            crosses.append(tc)
            curves.append(curve)
            print(f"# ttr {name} run {r}: {tt if tt is None else round(tt, 3)} s at {spp} spp "
                  f"(interpolated crossing {tc if tc is None else round(tc, 3)} s); last relMSE {curve[-1][2]:.3e}",
                  file=sys.stderr, flush=True)
        ok = [x for x in times if x is not None]
        okc = [x for x in crosses if x is not None]
        res[name] = {"seconds_median": float(np.median(ok)) if len(ok) == len(times) else None,
                     "crossing_seconds_median": float(np.median(okc)) if len(okc) == len(crosses) else None,
                     "seconds_runs": times, "crossing_seconds_runs": crosses, "curve_run0": curves[0]}
    sp = spc = None
    if res["plain"]["seconds_median"] and res["guided"]["seconds_median"]:
        sp = res["plain"]["seconds_median"] / res["guided"]["seconds_median"]
    if res["plain"]["crossing_seconds_median"] and res["guided"]["crossing_seconds_median"]:
        spc = res["plain"]["crossing_seconds_median"] / res["guided"]["crossing_seconds_median"]
    return {"target_relmse": a.ttr_target, "workload": f"guide_room seed 4 (1,003,550 tris, gap 0.2, L 500), "
                                                       f"{a.ttr_width}x{a.ttr_height}, max depth 16",
            "reference": {"spp": a.ttr_ref_spp, "estimator": "plain", "noise_floor_relmse": floor,
                          "seconds": round(ref_s, 2)},
            "guided_schedule": {"train_spp": a.ttr_train_spp, "spp_per_pass": 8, "adam_steps_per_pass": 32,
                                "batch": 1 << 18, "lr": 1e-2, "guide_fraction": 0.5},
            "plain": res["plain"], "guided": res["guided"], "speedup_guided_vs_plain": sp,
            "speedup_guided_vs_plain_interpolated": spc}


def medium_fp8(wr, scenes, dev, a):
    """The gap-lit room filled with a homogeneous medium (sigma_s 0.4, sigma_a 0.05
    RGB-equal, HG g 0.7), max depth 32, med_spp samples per pixel: plain NEE+MIS,
    and guided with BF16 and with FP8 (e4m3) network inference -- each guided run
    trains its own network online (BF16) for the first med_train_spp samples
    (8-spp passes, 32 Adam steps after each) and renders the rest with it.
    relMSE against a plain render at med_ref_spp (two halves, whose difference
    estimates the reference's own noise floor)."""
    N = 34
    sc = scenes.guide_room(4, resolution=(a.med_width, a.med_height)).with_medium(sigma_a=0.05, sigma_s=0.4, g=0.7)
    gs = wr.GpuScene(dev, sc)
    t = time.perf_counter()
    half = a.med_ref_spp // 2
    fa, fb = wr.Film(gs), wr.Film(gs)
    wr.render_nee(gs, fa, half, num_path_vertices=N, seed=7001)
    wr.render_nee(gs, fb, half, num_path_vertices=N, seed=7002)
    ia, ib = fa.image(), fb.image()
    ref = 0.5 * (ia + ib)
    ref_s = time.perf_counter() - t
    floor = rel_mse(ia, ib) / 4.0
    print(f"# medium reference: {a.med_ref_spp} spp in {ref_s:.1f}s, noise floor ~{floor:.2e}", file=sys.stderr,
          flush=True)
    res = {}
    film = wr.Film(gs)
    t = time.perf_counter()
    st = wr.render_nee(gs, film, a.med_spp, num_path_vertices=N, seed=7100)
    sec = time.perf_counter() - t
    res["plain"] = {"seconds": round(sec, 3), "relmse": rel_mse(film.image(), ref),
                    "mrays_per_s": round((st.extend_rays + st.shadow_rays) / sec / 1e6, 1)}
    for name, fp8 in (("guided_bf16", False), ("guided_fp8", True)):
        g = wr.Guide(gs, learning_rate=1e-2, batch_size=1 << 18, record_capacity=1 << 20, seed=7)
        film = wr.Film(gs)
        t = time.perf_counter()
        k = min(a.med_train_spp, a.med_spp)
        s1, _ = wr.render_guided(gs, film, g, k, num_path_vertices=N, seed=7200, guide_fraction=0.5, spp_per_pass=8,
                                 train_steps=32, warmup_passes=1, fp8_inference=fp8)
        rays = s1.extend_rays + s1.shadow_rays
        if a.med_spp > k:
            s2, _ = wr.render_guided(gs, film, g, a.med_spp - k, num_path_vertices=N, seed=7200, guide_fraction=0.5,
                                     spp_per_pass=a.med_spp - k, train_steps=0, warmup_passes=0, fp8_inference=fp8)
            rays += s2.extend_rays + s2.shadow_rays
        sec = time.perf_counter() - t
        res[name] = {"seconds": round(sec, 3), "relmse": rel_mse(film.image(), ref),
                     "mrays_per_s": round(rays / sec / 1e6, 1)}
        g.close()
        film.close()
    for k, v in res.items():
        print(f"# medium {k}: {v['seconds']} s, relMSE {v['relmse']:.4e}, {v['mrays_per_s']} Mrays/s", file=sys.stderr,
              flush=True)
    gs.close()
    return {"workload": f"guide_room seed 4 (1,003,550 tris, gap 0.2, L 500) in a homogeneous medium "
                        f"(sigma_s 0.4, sigma_a 0.05, HG g 0.7), {a.med_width}x{a.med_height}, {a.med_spp} spp, "
                        f"max depth 32",
            "reference": {"spp": a.med_ref_spp, "estimator": "plain GPU (pinned to the FP64 oracle in tests)",
                          "noise_floor_relmse": floor, "seconds": round(ref_s, 2)},
            "guided_schedule": {"train_spp": a.med_train_spp, "spp_per_pass": 8, "adam_steps_per_pass": 32,
                                "training": "bf16"},
            **res}


def batched(wr, scenes, dev, a):
    """The batched workload: 32 seeded Cornell-box scenes with icosphere clutter
    (10k..2M triangles, log-uniform), resolutions 128^2..1024^2, 16..256 spp, max
    depth 8..16, rendered as one batch (one launch of each kernel per iteration,
    scheduling from the autotuner), then one scene at a time for comparison.
    Per-scene relMSE is batched vs individual film (same random numbers)."""
    import numpy as np
    t0 = time.perf_counter()
    work = scenes.batch_workload(seed=3, n_scenes=a.batch_scenes)
    gsc = [wr.GpuScene(dev, sc) for sc, _, _ in work]
    build_s = time.perf_counter() - t0
    tris = sum(sc.n_tris for sc, _, _ in work)
    jobs = [dict(scene=g, spp=spp, num_path_vertices=N, seed=1000 + i)
            for i, (g, (sc, spp, N)) in enumerate(zip(gsc, work))]
    t0 = time.perf_counter()
    tuning, trial_rate = wr.autotune(jobs)
    tune_s = time.perf_counter() - t0
    for j in jobs:
        j["film"] = wr.Film(j["scene"])
    t0 = time.perf_counter()
    st = wr.render_batch(jobs, tuning=tuning)
    t_batch = time.perf_counter() - t0
    rays_b = st.extend_rays + st.shadow_rays
    print(f"# batch: {len(jobs)} scenes, {tris} tris, {st.paths} paths in {t_batch:.2f}s "
          f"({rays_b / t_batch / 1e6:.0f} Mrays/s); build {build_s:.1f}s, autotune {tune_s:.1f}s", file=sys.stderr,
          flush=True)
    t_ind, rays_i, rel = 0.0, 0, []
    for j in jobs:
        f = wr.Film(j["scene"])
        t0 = time.perf_counter()
        s1 = wr.render_nee(j["scene"], f, j["spp"], num_path_vertices=j["num_path_vertices"], seed=j["seed"])
        t_ind += time.perf_counter() - t0
        rays_i += s1.extend_rays + s1.shadow_rays
        rel.append(rel_mse(j["film"].image(), f.image()))
        f.close()
        j["film"].close()
        j["scene"].close()
    print(f"# individually: {t_ind:.2f}s ({rays_i / t_ind / 1e6:.0f} Mrays/s); max per-scene relMSE(batch, single) "
          f"{max(rel):.2e}", file=sys.stderr, flush=True)
    return {"workload": f"{len(jobs)} seeded Cornell boxes with icosphere clutter (scene i seed 1000+i), "
                        f"{tris} triangles, 128^2..1024^2, 16..256 spp, max depth 8..16",
            "paths": st.paths, "rays": rays_b, "seconds": round(t_batch, 3),
            "mrays_per_s": round(rays_b / t_batch / 1e6, 1),
            "individual_seconds": round(t_ind, 3), "individual_mrays_per_s": round(rays_i / t_ind / 1e6, 1),
            "speedup_vs_individual": round(t_ind / t_batch, 3),
            "autotuned": {"max_paths_in_flight": tuning.max_paths_in_flight, "fetch_chunk": tuning.fetch_chunk,
                          "refill_lanes": tuning.refill_lanes, "use_graph": bool(tuning.use_graph),
                          "seconds": round(tune_s, 2)},
            "per_scene_relmse_vs_individual": [float(f"{x:.3e}") for x in rel]}


def cpu_baseline(sc, N, target_s):
    """The oracle (FP64 restatement of the reference path) on the host cores:
    whole frames at 1 spp, successive sample indices, until ~target_s of work."""
    import oracle
    # every core this process may run on, capped by the CPU share the host grants
    # one GPU's job (OMP_NUM_THREADS where it is set)
    avail = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(avail, share) if share > 0 else avail
    nproc = os.cpu_count() or avail
    osc = oracle.OracleScene(sc)
    total_rays, total_s, total_paths, k = 0, 0.0, 0, 0
    while total_s < target_s:
        _, _, st = osc.render(1, num_path_vertices=N, seed=7, first_sample=k, threads=threads)
        total_rays += st.extend_rays + st.shadow_rays
        total_s += st.seconds
        total_paths += st.paths
        k += 1
    return {"value": round(total_rays / total_s / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 restatement, {k} full frame(s) at 1 spp = {total_paths} paths of the same "
                      f"scene and depth, {total_s:.1f} s on {threads} threads",
            "samples_per_s": round(total_paths / total_s, 1),
            "nproc": nproc, "affinity_cpus": avail, "cpu_share": share or None,
            "note": "the project's FP64 CPU restatement of the reference estimator (the reference integrator "
                    "needs Embree's CMake-generated headers and cannot be built here); threads = the CPUs this "
                    "process may use, capped by the host's CPU share for one GPU (OMP_NUM_THREADS)"}


def other_integrators(wr, scenes, dev, gs, sc, N, a):
    """The reference's other integrators on the same kernels, on the headline
    scene and frame: primitive_path_tracer and light_splatting at the headline
    depth (light_splatting traces spp * W * H light subpaths), visualize one
    primary hit per sample, NEE+MIS with a 1024 x 512 envmap sky added, and
    bidirectional_path_tracer at a quarter of their spp (every camera vertex
    connects to every light vertex).  Mrays/s = (closest-hit + shadow rays) /
    wall second of one blocking call."""
    import numpy as np
    out = {}
    spp = max(1, a.spp // 4)
    bspp = max(1, a.spp // 16)
    sky = np.ones((512, 1024, 3), np.float32) * np.float32([0.3, 0.4, 0.6])
    sky[40:48, 300:308] = 200.0  # a sun
    for name, n_spp, fn, scene in (
            ("primitive_path_tracer", spp, lambda g, f: wr.render_primitive(g, f, spp, num_path_vertices=N, seed=3), gs),
            ("light_splatting", spp, lambda g, f: wr.render_light(g, f, spp, num_path_vertices=N, seed=3), gs),
            ("visualize", spp, lambda g, f: wr.render_visualize(g, f, spp, seed=3), gs),
            ("nee_envmap_sky", spp, lambda g, f: wr.render_nee(g, f, spp, num_path_vertices=N, seed=3), None),
            ("bidirectional_path_tracer", bspp, lambda g, f: wr.render_bdpt(g, f, bspp, num_path_vertices=N, seed=3),
             gs)):
        g = scene
        if g is None:
            g = wr.GpuScene(dev, sc.with_envmap(sky))
        f = wr.Film(g)
        fn(g, f)  # warm-up (pool allocation, first launches)
        f.clear()
        st = fn(g, f)
        rays = st.extend_rays + st.shadow_rays
        out[name] = {"spp": n_spp, "paths": st.paths, "rays": rays, "ms": round(st.wall_ms, 2),
                     "mrays_per_s": round(rays / st.wall_ms / 1e3, 1)}
        print(f"# {name}: {n_spp} spp in {st.wall_ms:.1f} ms, {rays / st.wall_ms / 1e3:.0f} Mrays/s",
              file=sys.stderr, flush=True)
        f.close()
        if g is not gs:
            g.close()
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported in this build "
              f"(the path would shard samples across GPUs with one FP64 film reduction)", file=sys.stderr)
        sys.exit(2)
    import wurstrenderer_amd as wr
    from wurstrenderer_amd import scenes

    if a.calibrate_copy:
        calibrate_copy()
    N = a.depth + 2
    t0 = time.time()
    sc = scenes.icospheres(2, resolution=(a.width, a.height))
    dev = wr.Device(0)
    gs = wr.GpuScene(dev, sc)
    info = gs.info()
    film = wr.Film(gs)
    print(f"# scene {sc.name}: {sc.n_tris} tris, {info['nodes']} nodes, build+upload {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    # the BVH: GPU build (the default) against the host SAH build, invariants checked on both
    bvh = {}
    for name, g_ in (("gpu_ploc", gs), ("gpu_ploc_warm", None), ("host_sah", None)):
        if g_ is None:
            g_ = wr.GpuScene(dev, sc, builder="host" if name == "host_sah" else "gpu")
        st = g_.bvh_check(sc)  # also checks the tree's invariants
        bvh[name] = {"build_ms": round(st["build_ms"], 1), "nodes": st["n_nodes"], "depth": st["depth"],
                     "sah_cost": round(st["sah_cost"], 3)}
        if g_ is not gs:
            g_.close()
    bvh["note"] = "gpu_ploc: first build in the process (includes loading the build kernels); build_ms includes the mesh upload"
    print(f"# bvh: gpu {bvh['gpu_ploc']} host {bvh['host_sah']}", file=sys.stderr, flush=True)
    for i in range(a.warmup):
        wr.render_nee(gs, film, a.spp, num_path_vertices=N, seed=1)
    film.clear()
    # timed region: every render_nee call is blocking (it synchronises the scene's stream)
    stats = []
    t_start = time.perf_counter()
    for k in range(a.steps):
        stats.append(wr.render_nee(gs, film, a.spp, num_path_vertices=N, seed=1, profile=True))
    t_end = time.perf_counter()
    wall = t_end - t_start
    ext = sum(s.extend_rays for s in stats)
    shd = sum(s.shadow_rays for s in stats)
    paths = sum(s.paths for s in stats)
    mrays = (ext + shd) / wall / 1e6
    ext_ms = sum(s.extend_ms for s in stats)
    ext_launches = sum(s.extend_launches for s in stats)
    shd_ms = sum(s.shadow_ms for s in stats)
    print(f"# timed {a.steps} steps in {wall:.3f}s: ext {ext} shd {shd} rays; extend {ext_ms:.1f} ms "
          f"shadow {shd_ms:.1f} ms shade {sum(s.shade_ms for s in stats):.1f} ms "
          f"generate {sum(s.generate_ms for s in stats):.1f} ms", file=sys.stderr, flush=True)
    # roofline of the closest-hit traversal kernel (algorithmic bytes per launch / mean launch time)
    roof = None
    if not a.no_count:
        f2 = wr.Film(gs)
        cs = wr.render_nee(gs, f2, a.spp, num_path_vertices=N, seed=1, first_sample=0, count_traversal=True)
        nodes_per_ray = cs.extend_nodes / max(cs.extend_rays, 1)
        tris_per_ray = cs.extend_tris / max(cs.extend_rays, 1)
        bytes_total = ext * (EXT_RAY_BYTES + nodes_per_ray * NODE_BYTES + tris_per_ray * TRI_BYTES)
        per_launch_bytes = bytes_total / ext_launches
        avg_ms = ext_ms / ext_launches
        achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9
        traffic = None
        tp = os.path.join(ROOT, "profiles", "extend_traffic.json")
        if os.path.exists(tp):
            traffic = json.load(open(tp)).get("hbm_bytes_per_launch")
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                # algorithmic bytes count node reads the L2 / MALL serve, so frac can pass 1;
                # the DRAM-side rate of the same launches (rocprof counters):
                "traffic_gbs": None if traffic is None else round(traffic / (avg_ms * 1e-3) / 1e9, 1),
                "traffic_frac": None if traffic is None else round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "kernel": "k_extend_fast (closest-hit BVH4 traversal)",
                "avg_launch_ms": round(avg_ms, 4), "launches": ext_launches,
                "nodes_per_ray": round(nodes_per_ray, 2), "tris_per_ray": round(tris_per_ray, 2),
                "bytes_per_ray": round(EXT_RAY_BYTES + nodes_per_ray * NODE_BYTES + tris_per_ray * TRI_BYTES, 1)}
    others = None
    if not a.no_integrators:
        others = other_integrators(wr, scenes, dev, gs, sc, N, a)
    bat = None
    if not a.no_batch:
        bat = batched(wr, scenes, dev, a)
    ttr = None
    if not a.no_ttr:
        ttr = time_to_relmse(wr, scenes, dev, a)
    med = None
    if not a.no_medium:
        med = medium_fp8(wr, scenes, dev, a)
    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(sc, N, a.cpu_seconds)
    out = {
        "metric": "Mrays/s (plain GPU FP32 wavefront NEE+MIS, 1.31M-triangle scene)",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": "64 level-5 icospheres + floor (1,310,726 tris), 2 quad emitters",
                   "resolution": [a.width, a.height], "spp_per_step": a.spp, "max_depth": a.depth,
                   "num_path_vertices": N, "accumulation": "f64"},
        "samples_per_s": round(paths / wall, 1),
        "extend_rays": ext,
        "shadow_rays": shd,
        "kernel_ms": {k: round(sum(getattr(s, k) for s in stats), 2) for k in
                      ("extend_ms", "shadow_ms", "shade_ms", "generate_ms", "accumulate_ms")},
        "roofline": roof,
        "cpu_baseline": cpu,
        "time_to_relmse": ttr,
        "batched": bat,
        "medium_bf16_fp8": med,
        "bvh_build": bvh,
        "other_integrators": others,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
