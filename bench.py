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
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, published
NODE_BYTES, TRI_BYTES = 64, 48  # a traversal node (NodeQ) and a triangle record
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
    ap.add_argument("--no-ttr", action="store_true", help="(the default) no 480x270 time-to-relMSE measurement")
    ap.add_argument("--small-ttr", action="store_true",
                    help="also measure time-to-relMSE on the 480x270 gap room against a 65,536-spp reference")
    ap.add_argument("--ttr-width", type=int, default=480)
    ap.add_argument("--ttr-height", type=int, default=270)
    ap.add_argument("--ttr-target", type=float, default=1e-3)
    ap.add_argument("--ttr-ref-spp", type=int, default=65536)
    ap.add_argument("--ttr-runs", type=int, default=3)
    ap.add_argument("--ttr-train-spp", type=int, default=16)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-scene measurement")
    ap.add_argument("--no-integrators", action="store_true",
                    help="skip primitive / light tracing / visualize / envmap on the headline frame")
    ap.add_argument("--batch-scenes", type=int, default=32)
    ap.add_argument("--no-config4", action="store_true", help="skip the 1080p guided-vs-plain equal-relMSE measurement")
    ap.add_argument("--c4-width", type=int, default=1920)
    ap.add_argument("--c4-height", type=int, default=1080)
    ap.add_argument("--c4-ref-spp", type=int, default=8192)
    ap.add_argument("--c4-runs", type=int, default=3)
    ap.add_argument("--c4-plain-runs", type=int, default=2)
    ap.add_argument("--no-medium", action="store_true", help="skip the medium BF16/FP8 inference measurement")
    ap.add_argument("--med-width", type=int, default=1280)
    ap.add_argument("--med-height", type=int, default=720)
    ap.add_argument("--med-spp", type=int, default=512)
    ap.add_argument("--med-ref-spp", type=int, default=6144)
    ap.add_argument("--med-train-spp", type=int, default=16)
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


def checkpoints(first, max_spp, growth=1.25, step=8):
    """spp checkpoints: every `step` samples up to 128 (so online training runs
    whole passes), then x growth, rounded up to a multiple of `step`."""
    out, spp = [], first
    while spp <= max_spp:
        out.append(spp)
        nxt = spp + step if spp < 128 else int(math.ceil(spp * growth / step) * step)
        spp = max(nxt, spp + step)
    return out


def crossing(curve, target):
    """Wall time at which relMSE reaches target, interpolated in log-log between
    the two checkpoints around it (relMSE ~ 1/time there); None if never."""
    for k, (_, t1, e1) in enumerate(curve):
        if e1 <= target:
            if k == 0:
                return t1
            _, t0, e0 = curve[k - 1]
            if e0 > e1 > 0:
                return math.exp(math.log(t0) + (math.log(t1) - math.log(t0)) * math.log(e0 / target) /
                                math.log(e0 / e1))
            return t1
    return None


def progressive(render, film, ref, floor, target, points, done=0):
    """Renders up to each checkpoint in turn (only the render calls are timed) and
    records relMSE against ref minus the reference's own noise floor (for
    independent unbiased images E[(I-R)^2] = Var I + Var R); stops at the first
    checkpoint at or below target.  `done`: samples per pixel the film already
    holds.  Returns (seconds, crossing seconds, spp, curve)."""
    t_total, curve = 0.0, []
    for spp in points:
        t = time.perf_counter()
        render(film, spp - done)
        t_total += time.perf_counter() - t
        done = spp
        e = rel_mse(film.image(), ref) - floor
        curve.append([spp, round(t_total, 4), e])
        if e <= target:
            return t_total, crossing(curve, target), spp, curve
    return None, None, done, curve


def reference(wr, gs, spp, N, seeds):
    """Plain render of spp samples per pixel as two independent halves; returns
    (image, noise floor = relMSE of the mean of the halves, seconds)."""
    t = time.perf_counter()
    fa, fb = wr.Film(gs), wr.Film(gs)
    wr.render_nee(gs, fa, spp // 2, num_path_vertices=N, seed=seeds[0])
    wr.render_nee(gs, fb, spp // 2, num_path_vertices=N, seed=seeds[1])
    ia, ib = fa.image(), fb.image()
    fa.close()
    fb.close()
    return 0.5 * (ia + ib), rel_mse(ia, ib) / 4.0, time.perf_counter() - t


# The guided schedule every guided measurement uses: the network guides the
# first 5 bounces with probability 0.4 (deeper vertices sample the BSDF only)
# and trains online on the first 16 spp, as 2-spp passes each followed by 256
# Adam steps (lr 1e-2, batches of 2^18) over that pass's records (up to 2^22,
# a random subset of the pass's guided vertices), then renders the rest with
# the trained network.  Picked on the 1080p gap room at 256 spp (guided
# relMSE x time against plain's, two runs each): 8 x 8 spp with 64 steps
# 2.12x; 16 x 2 spp with 64 steps over 2^20 records 2.61-2.91x; 8 x 2 spp
# with 256 steps over 2^22 records 2.88-3.01x (192 steps 2.71-3.03x; 4 x 4 spp
# with 512 steps over 2^23 records 2.93-3.08x; 12 x 2 spp 2.56-2.90x).  A
# pass's record count is capped, so small passes give a full record set at a
# fraction of the render time, and the steps, not the samples, train the net.
GUIDE = dict(guide_fraction=0.4, max_guided_bounce=5, train_spp=16, spp_per_pass=2, train_steps=256,
             record_capacity=1 << 22)


def guided_renderer(wr, gs, N, seed, train_spp=GUIDE["train_spp"], fp8=False):
    """Online guided rendering as benchmarked (GUIDE): a fresh network, trained
    between the passes of the first train_spp samples per pixel, then used
    without further training."""
    g = wr.Guide(gs, learning_rate=1e-2, batch_size=1 << 18, record_capacity=GUIDE["record_capacity"], seed=seed)
    state = {"calls": 0}
    kw = dict(num_path_vertices=N, seed=seed, guide_fraction=GUIDE["guide_fraction"],
              max_guided_bounce=GUIDE["max_guided_bounce"], fp8_inference=fp8)

    def render(film, spp):
        left = max(0, train_spp - film.spp)
        if left:
            k = min(spp, left)
            wr.render_guided(gs, film, g, k, spp_per_pass=GUIDE["spp_per_pass"], train_steps=GUIDE["train_steps"],
                             warmup_passes=1 if state["calls"] == 0 else 0, **kw)
            spp -= k
        if spp:
            wr.render_guided(gs, film, g, spp, spp_per_pass=spp, train_steps=0, warmup_passes=0, **kw)
        state["calls"] += 1
    return render, g


def time_to_relmse(wr, scenes, dev, a):
    """Guided vs plain GPU NEE+MIS on the gap-lit room (1,003,550 triangles, light
    L = 500 hidden behind a 0.2 x 0.2 gap), max depth 16: wall time until the
    progressively rendered image reaches relMSE <= target (reference noise floor
    subtracted) against a plain reference at ttr_ref_spp; checkpoints every
    8 spp to 128, then x1.25; median of ttr_runs runs, each from an untrained
    network."""
    import numpy as np
    N = 18
    sc = scenes.guide_room(4, resolution=(a.ttr_width, a.ttr_height))
    gs = wr.GpuScene(dev, sc)
    ref, floor, ref_s = reference(wr, gs, a.ttr_ref_spp, N, (9001, 9002))
    print(f"# ttr reference: {a.ttr_ref_spp} spp in {ref_s:.1f}s, noise floor ~{floor:.2e}", file=sys.stderr, flush=True)
    points = checkpoints(16, a.ttr_ref_spp // 2)
    res = {}
    for name in ("plain", "guided"):
        times, crosses, curves = [], [], []
        for r in range(a.ttr_runs):
            film = wr.Film(gs)
            if name == "plain":
                render = (lambda seed: lambda f, spp: wr.render_nee(gs, f, spp, num_path_vertices=N, seed=seed))(100 + r)
                g = None
            else:
                render, g = guided_renderer(wr, gs, N, 100 + r, a.ttr_train_spp)
            tt, tc, spp, curve = progressive(render, film, ref, floor, a.ttr_target, points)
            film.close()
            if g is not None:
                g.close()
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
                     "seconds_runs": times, "crossing_seconds_runs": crosses,
                     "curve_run0": [[c[0], c[1], float(f"{c[2]:.4e}")] for c in curves[0]]}
    sp = spc = None
    if res["plain"]["seconds_median"] and res["guided"]["seconds_median"]:
        sp = res["plain"]["seconds_median"] / res["guided"]["seconds_median"]
    if res["plain"]["crossing_seconds_median"] and res["guided"]["crossing_seconds_median"]:
        spc = res["plain"]["crossing_seconds_median"] / res["guided"]["crossing_seconds_median"]
    gs.close()
    return {"target_relmse": a.ttr_target, "workload": f"guide_room seed 4 (1,003,550 tris, gap 0.2, L 500), "
                                                       f"{a.ttr_width}x{a.ttr_height}, max depth 16",
            "reference": {"spp": a.ttr_ref_spp, "estimator": "plain GPU (pinned to the FP64 oracle in tests)",
                          "noise_floor_relmse": floor, "seconds": round(ref_s, 2),
                          "note": "relMSE values have the reference's noise floor subtracted"},
            "checkpoints": "every 8 spp to 128, then x1.25",
            "guided_schedule": {**GUIDE, "train_spp": a.ttr_train_spp, "batch": 1 << 18, "lr": 1e-2},
            "plain": res["plain"], "guided": res["guided"], "speedup_guided_vs_plain": sp,
            "speedup_guided_vs_plain_interpolated": spc}


def config4(wr, scenes, dev, a):
    """The learned-guiding workload at its own size: the gap-lit room at
    1920x1080, max depth 16, against a plain reference of c4_ref_spp (two
    independent halves; relMSE values have its noise floor subtracted).
    Guided (the GUIDE schedule: online training over the first 16 spp) renders
    256 spp, then continues at x1.25 checkpoints until relMSE <= ttr_target;
    plain renders progressively from 64 spp at x1.25 checkpoints until it
    reaches ttr_target.  Reported: plain's time to the guided 256-spp image's
    relMSE (equal relMSE) and both estimators' times to ttr_target, crossings
    interpolated in log-log between checkpoints; medians over c4_runs guided
    and c4_plain_runs plain runs."""
    import numpy as np
    N = 18
    sc = scenes.guide_room(4, resolution=(a.c4_width, a.c4_height))
    gs = wr.GpuScene(dev, sc)
    ref, floor, ref_s = reference(wr, gs, a.c4_ref_spp, N, (8001, 8002))
    print(f"# config-4 reference: {a.c4_ref_spp} spp in {ref_s:.1f}s, noise floor ~{floor:.2e}", file=sys.stderr,
          flush=True)
    g_t, g_e, g_x, g_curves = [], [], [], []
    for r in range(a.c4_runs):
        render, g = guided_renderer(wr, gs, N, 300 + r)
        film = wr.Film(gs)
        t = time.perf_counter()
        render(film, GUIDE["train_spp"])
        render(film, 256 - GUIDE["train_spp"])
        g_t.append(time.perf_counter() - t)
        g_e.append(rel_mse(film.image(), ref) - floor)
        curve = [[256, round(g_t[-1], 4), g_e[-1]]]
        if g_e[-1] > a.ttr_target:
            _, _, _, more = progressive(render, film, ref, floor, a.ttr_target, checkpoints(320, a.c4_ref_spp, step=64),
                                        done=256)
            curve += [[c[0], round(c[1] + g_t[-1], 4), c[2]] for c in more]
        g_x.append(crossing(curve, a.ttr_target))
        g_curves.append(curve)
        film.close()
        g.close()
        print(f"# config-4 guided run {r}: 256 spp in {g_t[-1]:.3f} s, relMSE {g_e[-1]:.4e}; relMSE "
              f"{a.ttr_target:g} at {g_x[-1] if g_x[-1] is None else round(g_x[-1], 3)} s ({curve[-1][0]} spp)",
              file=sys.stderr, flush=True)
    target = float(np.median(g_e))
    points = checkpoints(64, a.c4_ref_spp, step=64)
    p_eq, p_x, p_curves = [], [], []
    for r in range(a.c4_plain_runs):
        film = wr.Film(gs)
        render = (lambda seed: lambda f, spp: wr.render_nee(gs, f, spp, num_path_vertices=N, seed=seed))(400 + r)
        _, _, spp, curve = progressive(render, film, ref, floor, a.ttr_target, points)
        film.close()
        p_eq.append(crossing(curve, target))
        p_x.append(crossing(curve, a.ttr_target))
        p_curves.append(curve)
        print(f"# config-4 plain run {r}: relMSE {target:.4e} at {p_eq[-1] if p_eq[-1] is None else round(p_eq[-1], 3)}"
              f" s, relMSE {a.ttr_target:g} at {p_x[-1] if p_x[-1] is None else round(p_x[-1], 3)} s ({spp} spp)",
              file=sys.stderr, flush=True)
    gs.close()
    med = lambda xs: float(np.median(xs)) if xs and all(x is not None for x in xs) else None
    tg, tp, xg, xp = float(np.median(g_t)), med(p_eq), med(g_x), med(p_x)
    return {"workload": f"guide_room seed 4 (1,003,550 tris, gap 0.2, L 500), {a.c4_width}x{a.c4_height}, max "
                        f"depth 16",
            "guided_schedule": {**GUIDE, "batch": 1 << 18, "lr": 1e-2, "parameter_average_decay": 0.99},
            "reference": {"spp": a.c4_ref_spp, "estimator": "plain GPU (pinned to the FP64 oracle in tests)",
                          "noise_floor_relmse": floor, "seconds": round(ref_s, 2),
                          "note": "relMSE values have the reference's noise floor subtracted"},
            "equal_relmse_at_guided_256spp": {
                "guided_seconds_median": round(tg, 3), "relmse_median": target, "guided_seconds_runs": g_t,
                "guided_relmse_runs": g_e, "plain_seconds_median": tp, "plain_seconds_runs": p_eq,
                "speedup_guided_vs_plain": None if tp is None else round(tp / tg, 3)},
            "time_to_relmse": {
                "target_relmse": a.ttr_target, "guided_seconds_median": xg, "guided_seconds_runs": g_x,
                "plain_seconds_median": xp, "plain_seconds_runs": p_x,
                "speedup_guided_vs_plain": None if xg is None or xp is None else round(xp / xg, 3),
                "guided_curve_run0": [[c[0], c[1], float(f"{c[2]:.4e}")] for c in (g_curves[0] if g_curves else [])],
                "plain_curve_run0": [[c[0], c[1], float(f"{c[2]:.4e}")] for c in (p_curves[0] if p_curves else [])]}}


def medium_fp8(wr, scenes, dev, a):
    """The medium workload: the seed-5 batched-workload scene (716,836
    triangles) filled with a homogeneous medium (sigma_s 0.4, sigma_a 0.05
    RGB-equal, HG g 0.7), 1280x720, max depth 32, med_spp samples per pixel.
    Plain NEE+MIS; one network trained online (BF16) over med_train_spp; then
    the same trained network renders med_spp with BF16 and with FP8 (e4m3)
    inference from the same random numbers (paired: their difference is the
    precision effect, not noise).  relMSE against a plain reference of
    med_ref_spp, noise floor subtracted."""
    import numpy as np
    N = 34
    sc = scenes.medium_clutter(5, resolution=(a.med_width, a.med_height))
    gs = wr.GpuScene(dev, sc)
    # reference: the guided estimator (unbiased; its mean is pinned to the plain
    # renderer's in the GPU tests) with its own network, as two independent halves
    t = time.perf_counter()
    halves = []
    for k in range(2):
        render, gref = guided_renderer(wr, gs, N, 7001 + k)
        f = wr.Film(gs)
        render(f, GUIDE["train_spp"])
        render(f, a.med_ref_spp // 2 - GUIDE["train_spp"])
        halves.append(f.image())
        f.close()
        gref.close()
    ref = 0.5 * (halves[0] + halves[1])
    floor = rel_mse(halves[0], halves[1]) / 4.0
    ref_s = time.perf_counter() - t
    print(f"# medium reference: {a.med_ref_spp} guided spp in {ref_s:.1f}s, noise floor ~{floor:.2e}", file=sys.stderr,
          flush=True)
    res = {}
    film = wr.Film(gs)
    t = time.perf_counter()
    st = wr.render_nee(gs, film, a.med_spp, num_path_vertices=N, seed=7100)
    sec = time.perf_counter() - t
    res["plain"] = {"seconds": round(sec, 3), "relmse": rel_mse(film.image(), ref) - floor,
                    "mrays_per_s": round((st.extend_rays + st.shadow_rays) / sec / 1e6, 1)}
    film.close()
    g = wr.Guide(gs, learning_rate=1e-2, batch_size=1 << 18, record_capacity=GUIDE["record_capacity"], seed=7)
    film = wr.Film(gs)
    t = time.perf_counter()
    wr.render_guided(gs, film, g, a.med_train_spp, num_path_vertices=N, seed=7150,
                     guide_fraction=GUIDE["guide_fraction"], max_guided_bounce=GUIDE["max_guided_bounce"],
                     spp_per_pass=GUIDE["spp_per_pass"], train_steps=GUIDE["train_steps"], warmup_passes=1)
    train_s = time.perf_counter() - t
    film.close()
    trained = g.get_params()
    imgs = {}
    for name, fp8 in (("guided_bf16", False), ("guided_fp8", True)):
        g.set_params(trained)
        film = wr.Film(gs)
        t = time.perf_counter()
        st, _ = wr.render_guided(gs, film, g, a.med_spp, num_path_vertices=N, seed=7200,
                                 guide_fraction=GUIDE["guide_fraction"], max_guided_bounce=GUIDE["max_guided_bounce"],
                                 spp_per_pass=a.med_spp, train_steps=0, warmup_passes=0, fp8_inference=fp8)
        sec = time.perf_counter() - t
        imgs[name] = film.image()
        res[name] = {"seconds": round(sec, 3), "relmse": rel_mse(imgs[name], ref) - floor,
                     "mrays_per_s": round((st.extend_rays + st.shadow_rays) / sec / 1e6, 1)}
        film.close()
    g.close()
    gs.close()
    paired = rel_mse(imgs["guided_fp8"], imgs["guided_bf16"])
    # the FP8 - BF16 relMSE gap per pixel, at common random numbers: its mean and
    # standard error over pixels (the reference's own noise enters both terms
    # alike and largely cancels)
    den = ref * ref + 1e-2
    dpx = (((imgs["guided_fp8"] - ref) ** 2 - (imgs["guided_bf16"] - ref) ** 2) / den).mean(-1).ravel()
    gap, gap_se = float(dpx.mean()), float(dpx.std() / np.sqrt(dpx.size))
    for k, v in res.items():
        print(f"# medium {k}: {v['seconds']} s, relMSE {v['relmse']:.4e}, {v['mrays_per_s']} Mrays/s", file=sys.stderr,
              flush=True)
    print(f"# medium: relMSE(fp8 image, bf16 image) at the same random numbers {paired:.3e}; relMSE gap fp8 - bf16 "
          f"{gap:.3e} +- {gap_se:.1e} (z = {gap / gap_se:.1f}); training {train_s:.2f} s", file=sys.stderr, flush=True)
    return {"workload": f"seed-5 batched-workload scene ({sc.n_tris:,} tris) in a homogeneous medium "
                        f"(sigma_s 0.4, sigma_a 0.05, HG g 0.7), {a.med_width}x{a.med_height}, {a.med_spp} spp, "
                        f"max depth 32",
            "reference": {"spp": a.med_ref_spp, "estimator": "guided GPU, two independent networks and halves "
                                                              "(unbiased: pinned to the plain mean in the GPU tests, "
                                                              "plain to the FP64 oracle)",
                          "noise_floor_relmse": floor, "seconds": round(ref_s, 2),
                          "note": "relMSE values have the reference's noise floor subtracted"},
            "guided_schedule": {**GUIDE, "train_spp": a.med_train_spp, "training": "bf16",
                                "training_seconds": round(train_s, 3),
                                "inference": "the same trained network, the same random numbers, BF16 vs FP8"},
            "fp8_vs_bf16_paired_relmse": paired,
            "fp8_minus_bf16_relmse": {"gap": gap, "standard_error": gap_se, "z": gap / gap_se,
                                      "note": "per-pixel difference of squared relative errors at common random "
                                              "numbers, mean and standard error over pixels"},
            **res}


def batched_small(wr, scenes, dev):
    """The same generator restricted to small jobs (32 scenes, 128^2 or 256^2,
    16..64 spp, 10k..200k triangles): single renders of these leave the GPU
    underfilled, which one batch does not.  Batch vs one scene at a time with
    the same autotuned scheduling, every scene warmed up first."""
    work = scenes.batch_workload(seed=3, n_scenes=32, resolutions=(128, 256), spp_range=(16, 64),
                                 tri_range=(10_000, 200_000))
    gsc = [wr.GpuScene(dev, sc) for sc, _, _ in work]
    jobs = [dict(scene=g, spp=spp, num_path_vertices=N, seed=2000 + i) for i, (g, (sc, spp, N)) in
            enumerate(zip(gsc, work))]
    tuning, _ = wr.autotune(jobs)
    for j in jobs:
        j["film"] = wr.Film(j["scene"])
    wr.render_batch(jobs, tuning=tuning)
    for j in jobs:
        j["film"].clear()
    t0 = time.perf_counter()
    st = wr.render_batch(jobs, tuning=tuning)
    t_batch = time.perf_counter() - t0
    t_ind, rel = 0.0, []
    for j in jobs:
        f = wr.Film(j["scene"])
        wr.render_nee(j["scene"], f, 1, num_path_vertices=j["num_path_vertices"], seed=j["seed"], tuning=tuning)
        f.clear()
        t0 = time.perf_counter()
        wr.render_nee(j["scene"], f, j["spp"], num_path_vertices=j["num_path_vertices"], seed=j["seed"], tuning=tuning)
        t_ind += time.perf_counter() - t0
        rel.append(rel_mse(j["film"].image(), f.image()))
        f.close()
        j["film"].close()
        j["scene"].close()
    rays = st.extend_rays + st.shadow_rays
    print(f"# small-job batch: {len(jobs)} scenes in {t_batch * 1e3:.1f} ms ({rays / t_batch / 1e6:.0f} Mrays/s), "
          f"individually {t_ind * 1e3:.1f} ms; max relMSE(batch, single) {max(rel):.2e}", file=sys.stderr, flush=True)
    return {"workload": "32 seeded Cornell boxes with icosphere clutter (scene i seed 2000+i), 128^2 or 256^2, "
                        "16..64 spp, 10k..200k triangles, max depth 8..16",
            "seconds": round(t_batch, 4), "individual_seconds": round(t_ind, 4),
            "mrays_per_s": round(rays / t_batch / 1e6, 1), "speedup_vs_individual": round(t_ind / t_batch, 3),
            "max_relmse_vs_individual": max(rel)}


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
    # the same batch with hipGraph replay switched the other way, so the line shows
    # what the autotuner's choice is worth; the films must agree with the first run
    alt = type(tuning).from_buffer_copy(tuning)
    alt.use_graph = 0 if tuning.use_graph else 1
    alt_jobs = [dict(j, film=wr.Film(j["scene"])) for j in jobs]
    t0 = time.perf_counter()
    wr.render_batch(alt_jobs, tuning=alt)
    t_alt = time.perf_counter() - t0
    alt_rel = max(rel_mse(j["film"].image(), aj["film"].image()) for j, aj in zip(jobs, alt_jobs))
    for aj in alt_jobs:
        aj["film"].close()
    print(f"# batch with use_graph={int(alt.use_graph)}: {t_alt:.2f}s; max per-scene relMSE against the first run "
          f"{alt_rel:.2e}", file=sys.stderr, flush=True)
    # one scene at a time, with the batch's scheduling parameters; each scene is
    # warmed up first (its path pool allocated, as the autotuner did for the batch)
    t_ind, rays_i, rel = 0.0, 0, []
    for j in jobs:
        f = wr.Film(j["scene"])
        wr.render_nee(j["scene"], f, 1, num_path_vertices=j["num_path_vertices"], seed=j["seed"], tuning=tuning)
        f.clear()
        t0 = time.perf_counter()
        s1 = wr.render_nee(j["scene"], f, j["spp"], num_path_vertices=j["num_path_vertices"], seed=j["seed"],
                           tuning=tuning)
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
            "graph_replay_switched": {"use_graph": bool(alt.use_graph), "seconds": round(t_alt, 3),
                                      "mrays_per_s": round(rays_b / t_alt / 1e6, 1),
                                      "max_relmse_vs_autotuned": float(f"{alt_rel:.3e}")},
            "per_scene_relmse_vs_individual": [float(f"{x:.3e}") for x in rel],
            "note": "one scene at a time uses the batch's autotuned scheduling and is warmed up first (pool "
                    "allocated), as the batch is by the autotuner; the large jobs fill the GPU by themselves, so "
                    "this mix gains little, the small-job batch below shows the batch where single renders cannot "
                    "fill the GPU",
            "small_jobs": batched_small(wr, scenes, dev)}


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
    print(f"# timed {a.steps} steps in {wall:.3f}s: ext {ext} shd {shd} rays in {ext_launches} "
          f"traversal launches ({ext_ms:.1f} ms) shade {sum(s.shade_ms for s in stats):.1f} ms "
          f"generate {sum(s.generate_ms for s in stats):.1f} ms", file=sys.stderr, flush=True)
    # roofline of the traversal kernel (k_trace_bounce_fast: one launch per pool
    # iteration traces that iteration's closest-hit rays and the previous one's
    # shadow rays): bytes leaving the L2 per launch (rocprof request counters of
    # this build on this workload, per traced ray, times the live rays per
    # launch) over the live mean launch time
    roof = None
    if not a.no_count:
        f2 = wr.Film(gs)
        cs = wr.render_nee(gs, f2, a.spp, num_path_vertices=N, seed=1, first_sample=0, count_traversal=True)
        f2.close()
        nodes_per_ray = cs.extend_nodes / max(cs.extend_rays, 1)
        tris_per_ray = cs.extend_tris / max(cs.extend_rays, 1)
        shd_nodes = cs.shadow_nodes / max(cs.shadow_rays, 1)
        shd_tris = cs.shadow_tris / max(cs.shadow_rays, 1)
        ext_bpr = EXT_RAY_BYTES + nodes_per_ray * NODE_BYTES + tris_per_ray * TRI_BYTES
        shd_bpr = SHD_RAY_BYTES + shd_nodes * NODE_BYTES + shd_tris * TRI_BYTES
        avg_ms = ext_ms / ext_launches
        ext_per_launch = ext / ext_launches
        shd_per_launch = shd / ext_launches
        rays_per_launch = ext_per_launch + shd_per_launch
        tp = os.path.join(ROOT, "profiles", "traversal_traffic.json")
        prof = json.load(open(tp)) if os.path.exists(tp) else {}
        tr_prof = prof.get("trace", {})
        roof = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None, "traffic": None,
                "kernel": "k_trace_bounce_fast (BVH4 traversal: closest-hit rays of a bounce with the shadow rays "
                          "of the one before)", "avg_launch_ms": round(avg_ms, 4),
                "launches": ext_launches, "rays_per_launch": round(rays_per_launch),
                "extend_rays_per_launch": round(ext_per_launch), "shadow_rays_per_launch": round(shd_per_launch),
                "traffic_source": "profiles/traversal_traffic.json (build " + str(prof.get("build")) + "): bytes leaving "
                                  "the L2 toward the fabric per traced ray (closest-hit and shadow), from rocprofv3 "
                                  "TCC_EA0_RDREQ_128B/64B/32B and WRITE_SIZE over the whole frame; the 256 MB MALL on "
                                  "that side serves part of them before HBM, so this bounds the HBM bytes from above"}
        if tr_prof.get("l2_egress_bytes_per_ray"):
            traffic = tr_prof["l2_egress_bytes_per_ray"] * rays_per_launch
            achieved = traffic / (avg_ms * 1e-3) / 1e9
            roof.update(achieved=round(achieved, 1), frac=round(achieved / HBM_PEAK_GBS, 4), traffic=round(traffic))
        # the cache-inclusive rate: every visited node and tested triangle counted
        # as read (the L2 and MALL serve most of them), so it can pass the HBM peak
        alg_bytes = ext_per_launch * ext_bpr + shd_per_launch * shd_bpr
        alg = alg_bytes / (avg_ms * 1e-3) / 1e9
        roof["algorithmic_incl_cache"] = {
            "bytes_per_launch": round(alg_bytes),
            "extend_bytes_per_ray": round(ext_bpr, 1), "nodes_per_ray": round(nodes_per_ray, 2),
            "tris_per_ray": round(tris_per_ray, 2), "shadow_bytes_per_ray": round(shd_bpr, 1),
            "shadow_nodes_per_ray": round(shd_nodes, 2), "shadow_tris_per_ray": round(shd_tris, 2),
            "gbs": round(alg, 1), "frac_of_hbm_peak": round(alg / HBM_PEAK_GBS, 4)}
        # the bound that traversal actually meets: the L1's rate of scattered lane
        # requests.  Lane loads per launch (ray in, four 16-byte rows per node
        # visit, three per triangle test, the result out; a shadow ray also reads
        # its contribution and target and updates its radiance) against the rate
        # that a full grid of independent scattered 16-byte loads reaches
        # (wr_debug_l1_request_probe: a 16 KiB buffer, which every L1 holds, and
        # a 4 MiB one, which the L2s hold and the L1s do not).
        ext_loads = 3 + 4 * nodes_per_ray + 3 * tris_per_ray
        shd_loads = 5 + 4 * shd_nodes + 3 * shd_tris
        lane_loads = ext_per_launch * ext_loads + shd_per_launch * shd_loads
        probes = {}
        for fp in (1 << 14, 1 << 22):
            p_ms, p_loads, n_cu = wr.l1_request_probe(dev, fp, 4096)
            probes[f"{fp >> 10}KiB"] = {"ms": round(p_ms, 3), "loads": p_loads,
                                        "per_ns_per_cu": round(p_loads / (p_ms * 1e6) / n_cu, 4)}
        trace_rate = lane_loads / (avg_ms * 1e6) / n_cu
        ceiling = probes["16KiB"]["per_ns_per_cu"]
        roof["l1_requests"] = {
            "lane_loads_per_launch": round(lane_loads), "extend_loads_per_ray": round(ext_loads, 2),
            "shadow_loads_per_ray": round(shd_loads, 2), "per_ns_per_cu": round(trace_rate, 4),
            "probe": probes, "frac_of_probe": round(trace_rate / ceiling, 4),
            "note": "the traversal kernel is bound by the rate of scattered lane requests at the L1, not by HBM "
                    "(with 128-byte nodes one extra 16-byte load per node visit cost ~7%, 30 extra VALU operations "
                    "~1%); the 16 KiB probe is the rate of independent scattered 16-byte loads that all hit the L1, "
                    "the 4 MiB one the rate when each must fetch its line from the L2; frac_of_probe counts the "
                    "kernel's lane loads, of which the L1 merges those of one instruction to one line"}
        # the compulsory rate: rays in and results out, plus the whole BVH read once per launch
        bvh_bytes = info["nodes"] * NODE_BYTES + sc.n_tris * TRI_BYTES
        comp_bytes = ext_per_launch * EXT_RAY_BYTES + shd_per_launch * SHD_RAY_BYTES + bvh_bytes
        comp = comp_bytes / (avg_ms * 1e-3) / 1e9
        roof["compulsory"] = {"bytes_per_launch": round(comp_bytes),
                              "gbs": round(comp, 1), "frac_of_hbm_peak": round(comp / HBM_PEAK_GBS, 4)}
    others = None
    if not a.no_integrators:
        others = other_integrators(wr, scenes, dev, gs, sc, N, a)
    bat = None
    if not a.no_batch:
        bat = batched(wr, scenes, dev, a)
    ttr = None
    if a.small_ttr:
        ttr = time_to_relmse(wr, scenes, dev, a)
    c4 = None
    if not a.no_config4:
        c4 = config4(wr, scenes, dev, a)
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
        # trace_ms: the traversal launches (closest-hit and shadow rays together)
        "kernel_ms": {name: round(sum(getattr(s, k) for s in stats), 2) for name, k in
                      (("trace_ms", "extend_ms"), ("shade_ms", "shade_ms"), ("generate_ms", "generate_ms"),
                       ("accumulate_ms", "accumulate_ms"))},
        "roofline": roof,
        "cpu_baseline": cpu,
        "guided_vs_plain_1080p": c4,
        "time_to_relmse_480x270": ttr,
        "batched": bat,
        "medium_bf16_fp8": med,
        "bvh_build": bvh,
        "other_integrators": others,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
