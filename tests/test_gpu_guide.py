"""GPU path-guiding network against the FP64 oracle (oracle/guide.py), and the
guided renderer against the plain one.

Bars, written per test:
  - forward: GPU (BF16 MFMA, FP32 accumulation) vs the oracle with the same BF16
    rounding points: relative L2 error < 2e-3, max abs error < 3e-2;
  - gradient: relative L2 error < 1e-2 per layer and cosine > 0.9999 overall
    (the oracle's own gradient is pinned by finite differences);
  - Adam: parameters after each of 3 steps within 2e-6 of the oracle's
    bias-corrected update fed with the GPU's gradient;
  - mixture: pdf rtol 1e-4; samples from the same uniforms within 2e-3 away
    from lobe-selection boundaries;
  - guided render: unbiased for any network -- the image mean within 4 sigma
    of the plain GPU render, which matches the oracle elsewhere;
  - gain: after training on the gap-lit room, guided relMSE at equal spp
    below the plain renderer's;
  - FP8 MFMA: v_mfma_f32_32x32x16_fp8_fp8 does not sum its exact e4m3
    products exactly: measured through a pass-through network, its error stays
    within 2^-13 of the sum of the products' magnitudes (about 14 significant
    bits; typical 3e-6);
  - FP8 forward, exact: on a signed generalised permutation network (one
    weight per neuron, so every MFMA sum is one exact product) with biases,
    the GPU equals the oracle bit for bit through all five layers -- per-layer
    weight scales, per-record activation scales, ReLU, operand layout;
  - FP8 forward, dense: with the accumulator inexact, hidden values near an
    e4m3 rounding midpoint may round the other way; 95% of records within
    1e-4 of the oracle and the relative L2 error under a fifth of the FP8
    quantisation error itself (FP8 oracle vs FP64);
  - FP8 render: unbiased for any network (as above); after BF16 training in a
    scattering medium, FP8 inference's relMSE within 1.3x of BF16 inference's
    at equal spp."""
import numpy as np
import pytest

from oracle import guide as G
import wurstrenderer_amd as wr
from wurstrenderer_amd import scenes

pytestmark = pytest.mark.gpu


def scene_box(sc):
    v = sc.verts.astype(np.float32)
    return np.concatenate([v.min(0), v.max(0)]).astype(np.float64)


def records(rng, n, box):
    lo, hi = box[:3], box[3:]
    pos = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)

    def unit(k):
        v = rng.normal(size=(k, 3))
        return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)

    return pos, unit(n), unit(n), unit(n), (rng.random(n) + 0.1).astype(np.float32)


def params(rng, scale_out=0.5):
    layers = []
    for l, (k, n) in enumerate(G.SHAPES):
        lim = np.sqrt(6.0 / k) * (scale_out if l == G.N_LAYERS - 1 else 1.0)
        layers.append((rng.uniform(-lim, lim, (k, n)), rng.uniform(-0.1, 0.1, n)))
    return G.pack(layers).astype(np.float32)


@pytest.fixture(scope="module")
def guide_scene(gpu_device):
    sc = scenes.cornell(1, resolution=(64, 64))
    return sc, wr.GpuScene(gpu_device, sc)


def test_forward_matches_oracle(guide_scene):
    sc, gs = guide_scene
    print("guide forward: 1000 records", flush=True)
    g = wr.Guide(gs)
    rng = np.random.default_rng(10)
    p = params(rng)
    g.set_params(p)
    np.testing.assert_array_equal(g.get_params(), p)
    box = scene_box(sc)
    pos, nrm, wo, _, _ = records(rng, 1000, box)
    out = g.forward(pos, nrm, wo)
    ref, _ = G.forward(p, pos, nrm, wo, box, bf=True)
    fp64, _ = G.forward(p, pos, nrm, wo, box)
    err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    err64 = np.linalg.norm(out - fp64) / np.linalg.norm(fp64)
    print(f"  rel err vs bf16 oracle {err:.2e}, vs fp64 oracle {err64:.2e}, max abs {np.abs(out - ref).max():.2e}")
    assert err < 2e-3 and np.abs(out - ref).max() < 3e-2
    # the GPU is closer to the BF16 emulation than to FP64 (it really rounds there)
    assert err < err64


FP8_ACC_REL = 2.0 ** -13


def _identity_layers(last):
    layers = []
    for l, (k, m) in enumerate(G.SHAPES):
        w = np.zeros((k, m))
        if l < G.N_LAYERS - 1:
            w[np.arange(min(k, m)), np.arange(min(k, m))] = 1.0
        else:
            w[:32] = last
        layers.append((w, np.zeros(m)))
    return G.pack(layers).astype(np.float32)


def test_fp8_mfma_accumulation_error_bound(guide_scene):
    """Layers 0-3 pass the e4m3 features through unchanged (exactly: identity
    weights are 128 I in e4m3), so the last layer's FP32 output, divided by its
    power-of-two scales, is the MFMA's raw sum of known exact products."""
    sc, gs = guide_scene
    print("fp8 mfma dot-product error: 2000 records x 40 outputs", flush=True)
    g = wr.Guide(gs)
    rng = np.random.default_rng(15)
    box = scene_box(sc)
    pos, nrm, wo, _, _ = records(rng, 2000, box)
    f = G.features(G.encode(pos, nrm, wo, box)).astype(np.float32).astype(np.float64)
    kx = G.pow2_exponent(f.astype(np.float32).max(axis=1))
    x8 = G.e4m3(np.ldexp(f, -kx[:, None]))
    worst = 0.0
    for w in (rng.uniform(-1, 1, (32, 40)), rng.uniform(0, 1, (32, 40))):
        p = _identity_layers(w)
        g.set_params(p)
        out = g.forward(pos, nrm, wo, fp8=True).astype(np.float64)
        np.testing.assert_array_equal(G.forward8(_identity_layers(np.eye(32, 40)), pos, nrm, wo, box)[:, :32],
                                      G.e4m3(np.ldexp(f, -kx[:, None])) * np.ldexp(1.0, kx)[:, None])
        kw = int(G.pow2_exponent(np.float32(np.abs(w).max())))
        w8 = G.e4m3(np.ldexp(w, -kw))
        acc = out / np.ldexp(1.0, kw + kx)[:, None]
        rel = np.abs(acc - x8 @ w8) / np.maximum(np.abs(x8) @ np.abs(w8), 1e-30)
        print(f"  |sum error| / sum |products|: max {rel.max():.2e} median {np.median(rel):.2e}")
        worst = max(worst, rel.max())
    assert worst <= FP8_ACC_REL


def test_fp8_forward_exact_on_permutation_network(guide_scene):
    sc, gs = guide_scene
    print("guide fp8 forward, signed permutation network: 1000 records", flush=True)
    g = wr.Guide(gs)
    rng = np.random.default_rng(16)
    vals = np.array([0.5, 0.75, 1.0, 1.25, 1.5])
    layers = []
    for k, n in G.SHAPES:
        w = np.zeros((k, n))
        w[rng.integers(0, k, n), np.arange(n)] = rng.choice(vals, n) * rng.choice([-1.0, 1.0, 1.0, 1.0], n)
        layers.append((w, rng.choice([0.0, 0.125, -0.25, 0.5], n)))
    p = G.pack(layers).astype(np.float32)
    g.set_params(p)
    box = scene_box(sc)
    pos, nrm, wo, _, _ = records(rng, 1000, box)
    out = g.forward(pos, nrm, wo, fp8=True)
    ref = G.forward8(p, pos, nrm, wo, box)
    print(f"  max |gpu - oracle| {np.abs(out - ref).max():.2e}, outputs nonzero {(ref != 0).mean():.2f}")
    assert (ref != 0).mean() > 0.3
    np.testing.assert_array_equal(out, ref.astype(np.float32))


def test_fp8_forward_matches_oracle(guide_scene):
    sc, gs = guide_scene
    print("guide fp8 forward, dense network: 1000 records", flush=True)
    g = wr.Guide(gs)
    rng = np.random.default_rng(14)
    p = params(rng)
    g.set_params(p)
    box = scene_box(sc)
    pos, nrm, wo, _, _ = records(rng, 1000, box)
    out = g.forward(pos, nrm, wo, fp8=True).astype(np.float64)
    ref8 = G.forward8(p, pos, nrm, wo, box)
    fp64, _ = G.forward(p, pos, nrm, wo, box)
    err = np.linalg.norm(out - ref8) / np.linalg.norm(ref8)
    quant = np.linalg.norm(ref8 - fp64) / np.linalg.norm(fp64)
    rec = np.abs(out - ref8).max(axis=1)
    print(f"  rel err vs fp8 oracle {err:.2e}; fp8 quantisation error {quant:.2e}; "
          f"records off by > 1e-4: {(rec > 1e-4).sum()}")
    assert (rec <= 1e-4).mean() >= 0.95
    assert err < 0.2 * quant


def test_gradient_matches_oracle(guide_scene):
    sc, gs = guide_scene
    print("guide gradient: 333 records (ragged last tile)", flush=True)
    g = wr.Guide(gs)
    rng = np.random.default_rng(11)
    p = params(rng, scale_out=1.0)
    g.set_params(p)
    box = scene_box(sc)
    pos, nrm, wo, wi, w = records(rng, 333, box)
    loss, grad = g.gradient(pos, nrm, wo, wi, w)
    rloss, rgrad = G.loss_and_grad(p, pos, nrm, wo, wi, w, box, bf=True)
    print(f"  loss gpu {loss:.6f} oracle {rloss:.6f}")
    assert abs(loss - rloss) < 1e-3 * abs(rloss)
    o = 0
    for l, (k, n) in enumerate(G.SHAPES):
        for name, sz in (("W", k * n), ("b", n)):
            a, b = grad[o:o + sz], rgrad[o:o + sz]
            e = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            print(f"  layer {l} {name}: rel err {e:.2e}")
            assert e < 1e-2, (l, name, e)
            o += sz
    cos = grad @ rgrad / (np.linalg.norm(grad) * np.linalg.norm(rgrad))
    assert cos > 0.9999


def test_adam_steps_match_oracle(guide_scene):
    """The Adam kernel against the bias-corrected update of the oracle, fed with
    the GPU's own gradient at each step (the gradient is pinned separately;
    Adam's per-element normalisation would amplify BF16 noise in elements with
    tiny gradients into +-lr differences)."""
    sc, gs = guide_scene
    print("guide adam: 3 steps on 256 records", flush=True)
    g = wr.Guide(gs, learning_rate=1e-2)
    rng = np.random.default_rng(12)
    p = params(rng)
    g.set_params(p)
    box = scene_box(sc)
    pos, nrm, wo, wi, w = records(rng, 256, box)
    q, m, v = p.astype(np.float64), np.zeros(G.NUM_PARAMS), np.zeros(G.NUM_PARAMS)
    for t in (1, 2, 3):
        _, grad = g.gradient(pos, nrm, wo, wi, w)
        g.train(pos, nrm, wo, wi, w, steps=1)
        q, m, v = G.adam_step(q, m, v, grad.astype(np.float64), t, lr=1e-2)
        got = g.get_params().astype(np.float64)
        err = np.abs(got - q).max()
        print(f"  step {t}: max |param - oracle| {err:.2e}, max step {np.abs(got - p).max():.2e}")
        assert err < 2e-6
    assert (np.abs(g.get_params() - p) > 0).mean() > 0.5


def test_mixture_matches_oracle(guide_scene):
    _, gs = guide_scene
    print("guide mixture: 4096 rows", flush=True)
    g = wr.Guide(gs)
    rng = np.random.default_rng(13)
    n = 4096
    raw = rng.normal(size=(n, 40)).astype(np.float32)
    raw[:, 4::5] = rng.uniform(-2, 6, size=(n, 8))  # sharpness 0.13 .. 6
    raw[::7, 4::5] = rng.uniform(20, 200, size=(len(raw[::7]), 8))  # and very sharp lobes
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    u = rng.random((n, 3)).astype(np.float32)
    pd, smp, sp = g.mixture(raw, d, u)
    np.testing.assert_allclose(pd, G.pdf(raw, d), rtol=1e-4, atol=1e-30)
    ref, _ = G.sample(raw, u)
    logw, _, _ = G.decode(raw)
    c = np.cumsum(np.exp(logw), axis=1)
    far = np.abs(u[:, :1] - c[:, :-1]).min(axis=1) > 1e-4
    err = np.linalg.norm(smp - ref, axis=1)
    print(f"  sample dir err: median {np.median(err):.1e}, max away from boundaries {err[far].max():.1e}")
    assert far.mean() > 0.95 and err[far].max() < 2e-3
    np.testing.assert_allclose(sp, G.pdf(raw, smp.astype(np.float64)), rtol=2e-3)


def _render(gs, spp, N, seed, guide=None, **kw):
    film = wr.Film(gs)
    if guide is None:
        st = wr.render_nee(gs, film, spp, num_path_vertices=N, seed=seed)
        return film.image(), st, None
    st, gst = wr.render_guided(gs, film, guide, spp, num_path_vertices=N, seed=seed, **kw)
    return film.image(), st, gst


@pytest.mark.parametrize("fp8", [False, True], ids=["bf16", "fp8"])
def test_guided_render_is_unbiased_with_any_network(gpu_device, fp8):
    """A random network mixed in with alpha = 0.5 changes the noise, not the mean."""
    sc = scenes.cornell(1, resolution=(48, 48))
    gs = wr.GpuScene(gpu_device, sc)
    print(f"guided render, random network ({'fp8' if fp8 else 'bf16'}): cornell 48x48", flush=True)
    g = wr.Guide(gs, seed=5)
    spp = 1024
    plain, _, _ = _render(gs, spp, 10, seed=21)
    guided, st, gst = _render(gs, spp, 10, seed=22, guide=g, guide_fraction=0.5, spp_per_pass=spp, train_steps=0,
                              warmup_passes=0, fp8_inference=fp8)
    assert gst.records == 0 and gst.train_steps == 0
    assert np.isfinite(guided).all() and st.extend_rays > 0
    # per-pixel noise of the two estimators from two more independent renders
    plain2, _, _ = _render(gs, spp, 10, seed=23)
    guided2, _, _ = _render(gs, spp, 10, seed=24, guide=g, guide_fraction=0.5, spp_per_pass=spp, train_steps=0,
                            warmup_passes=0, fp8_inference=fp8)
    npx = guided.shape[0] * guided.shape[1]
    sig = np.sqrt(((guided - guided2) ** 2).mean((0, 1)) / 2 / npx + ((plain - plain2) ** 2).mean((0, 1)) / 2 / npx)
    diff = guided.mean((0, 1)) - plain.mean((0, 1))
    print(f"  mean plain {plain.mean((0, 1))} guided {guided.mean((0, 1))} sigma {sig}")
    assert np.all(np.abs(diff) < 4 * sig + 1e-6), (diff, sig)


@pytest.mark.parametrize("max_bounce", [1, 3])
def test_guided_render_with_bounce_limit_is_unbiased(gpu_device, max_bounce):
    """Guiding only the first max_bounce vertices (deeper ones sample the BSDF
    with probability 1) changes the noise, not the mean; deeper vertices take
    no network draw, so they record no training data either."""
    sc = scenes.cornell(1, resolution=(48, 48))
    gs = wr.GpuScene(gpu_device, sc)
    print(f"guided render, random network, first {max_bounce} bounces: cornell 48x48", flush=True)
    g = wr.Guide(gs, seed=5)
    spp = 1024
    kw = dict(guide_fraction=0.5, spp_per_pass=spp, train_steps=0, warmup_passes=0, max_guided_bounce=max_bounce)
    plain, _, _ = _render(gs, spp, 10, seed=25)
    plain2, _, _ = _render(gs, spp, 10, seed=26)
    guided, st, _ = _render(gs, spp, 10, seed=27, guide=g, **kw)
    guided2, _, _ = _render(gs, spp, 10, seed=28, guide=g, **kw)
    assert np.isfinite(guided).all() and st.extend_rays > 0
    npx = guided.shape[0] * guided.shape[1]
    sig = np.sqrt(((guided - guided2) ** 2).mean((0, 1)) / 2 / npx + ((plain - plain2) ** 2).mean((0, 1)) / 2 / npx)
    diff = guided.mean((0, 1)) - plain.mean((0, 1))
    print(f"  mean plain {plain.mean((0, 1))} guided {guided.mean((0, 1))} sigma {sig}")
    assert np.all(np.abs(diff) < 4 * sig + 1e-6), (diff, sig)
    # the limit takes effect: at the same random numbers the image differs from
    # an unlimited guide's (deeper vertices draw from the BSDF instead)
    if max_bounce == 1:
        full, _, _ = _render(gs, spp, 10, seed=27, guide=g, **dict(kw, max_guided_bounce=0))
        assert np.abs(guided - full).max() > 0


def test_guided_render_trains_and_stays_unbiased(gpu_device):
    sc = scenes.guide_room(4, n_spheres=6, level=3, resolution=(64, 48))
    gs = wr.GpuScene(gpu_device, sc)
    print("guided render with training: guide room 64x48", flush=True)
    g = wr.Guide(gs, batch_size=1 << 14, record_capacity=1 << 16)
    before = g.get_params()
    img, st, gst = _render(gs, 256, 18, seed=31, guide=g, guide_fraction=0.5, spp_per_pass=32, train_steps=8,
                           warmup_passes=1)
    print(f"  records {gst.records} steps {gst.train_steps} loss {gst.last_loss:.4f} train_ms {gst.train_ms:.1f}")
    assert gst.records > 1000 and gst.train_steps == 8 * 8 and np.isfinite(gst.last_loss)
    assert np.isfinite(img).all()
    assert np.abs(g.get_params() - before).max() > 0
    ref, _, _ = _render(gs, 2048, 18, seed=32)
    plain, _, _ = _render(gs, 256, 18, seed=33)
    plain2, _, _ = _render(gs, 256, 18, seed=34)
    npx = img.shape[0] * img.shape[1]
    sig = np.sqrt(2 * ((plain - plain2) ** 2).mean((0, 1)) / 2 / npx)
    print(f"  mean ref {ref.mean((0, 1))} guided {img.mean((0, 1))} sigma {sig}")
    assert np.all(np.abs(img.mean((0, 1)) - ref.mean((0, 1))) < 5 * sig + 1e-6)


@pytest.mark.parametrize("overlap,avg", [(True, None), (False, 0.0)], ids=["overlap", "no-average"])
def test_overlapped_training_and_parameter_average_stay_unbiased(gpu_device, overlap, avg):
    """overlap_training renders pass q with the network trained on pass q - 2
    while pass q - 1 trains on its own stream; the parameter average (avg_decay,
    0 = the parameters themselves) decides which network the renderer samples.
    Neither changes the estimator's mean, and both train the network."""
    sc = scenes.guide_room(4, n_spheres=6, level=3, resolution=(64, 48))
    gs = wr.GpuScene(gpu_device, sc)
    print(f"guided render, overlap {overlap}, avg_decay {avg}: guide room 64x48", flush=True)
    g = wr.Guide(gs, batch_size=1 << 14, record_capacity=1 << 16, avg_decay=avg)
    before = g.get_params()
    img, st, gst = _render(gs, 256, 18, seed=35, guide=g, guide_fraction=0.4, spp_per_pass=16, train_steps=16,
                           warmup_passes=2 if overlap else 1, overlap_training=overlap, max_guided_bounce=5)
    print(f"  records {gst.records} steps {gst.train_steps} loss {gst.last_loss:.4f} train_ms {gst.train_ms:.1f}")
    assert gst.train_steps == 16 * 16 and np.isfinite(gst.last_loss) and np.abs(g.get_params() - before).max() > 0
    assert np.isfinite(img).all() and img.min() >= 0.0
    ref, _, _ = _render(gs, 2048, 18, seed=36)
    plain, _, _ = _render(gs, 256, 18, seed=37)
    plain2, _, _ = _render(gs, 256, 18, seed=38)
    npx = img.shape[0] * img.shape[1]
    sig = np.sqrt(2 * ((plain - plain2) ** 2).mean((0, 1)) / 2 / npx)
    print(f"  mean ref {ref.mean((0, 1))} guided {img.mean((0, 1))} sigma {sig}")
    assert np.all(np.abs(img.mean((0, 1)) - ref.mean((0, 1))) < 5 * sig + 1e-6)


def test_trained_guiding_beats_plain_at_equal_spp(gpu_device):
    sc = scenes.guide_room(4, n_spheres=12, level=3, resolution=(96, 64))
    gs = wr.GpuScene(gpu_device, sc)
    print("guided vs plain at equal spp: guide room 96x64", flush=True)
    ref, _, _ = _render(gs, 16384, 18, seed=41)
    g = wr.Guide(gs, batch_size=1 << 15, record_capacity=1 << 17)
    # train, then render with the trained network
    _render(gs, 256, 18, seed=42, guide=g, guide_fraction=0.5, spp_per_pass=16, train_steps=8, warmup_passes=1)
    guided, _, _ = _render(gs, 256, 18, seed=43, guide=g, guide_fraction=0.5, spp_per_pass=256, train_steps=0,
                           warmup_passes=0)
    plain, _, _ = _render(gs, 256, 18, seed=44)
    import oracle
    e_g, e_p = oracle.rel_mse(guided, ref), oracle.rel_mse(plain, ref)
    print(f"  relMSE at 256 spp: plain {e_p:.4e} guided {e_g:.4e} (gain {e_p / e_g:.2f}x)")
    assert e_g < e_p


def test_fp8_inference_in_medium_matches_bf16(gpu_device):
    """Config-5 shape at small size: clutter in a scattering medium (sigma_s 0.4,
    sigma_a 0.05, HG g 0.7), network trained in BF16, then rendered with BF16
    and with FP8 inference: both unbiased against a plain reference, FP8 noise
    close to BF16's."""
    sc = scenes.cornell_clutter(5, 20_000, resolution=(64, 48)).with_medium(sigma_a=0.05, sigma_s=0.4, g=0.7)
    gs = wr.GpuScene(gpu_device, sc)
    print("fp8 vs bf16 inference in a medium: clutter 20k 64x48", flush=True)
    ref, _, _ = _render(gs, 8192, 34, seed=51)
    g = wr.Guide(gs, batch_size=1 << 15, record_capacity=1 << 17)
    _render(gs, 256, 34, seed=52, guide=g, guide_fraction=0.5, spp_per_pass=16, train_steps=8, warmup_passes=1)
    import oracle
    res = {}
    for fp8, seeds in ((False, (53, 54)), (True, (55, 56))):
        imgs = [_render(gs, 256, 34, seed=sd, guide=g, guide_fraction=0.5, spp_per_pass=256, train_steps=0,
                        warmup_passes=0, fp8_inference=fp8)[0] for sd in seeds]
        res[fp8] = imgs
    e_bf, e_f8 = oracle.rel_mse(res[False][0], ref), oracle.rel_mse(res[True][0], ref)
    print(f"  relMSE at 256 spp vs 8192-spp plain: bf16 {e_bf:.4e} fp8 {e_f8:.4e}")
    assert e_f8 < 1.3 * e_bf
    npx = ref.shape[0] * ref.shape[1]
    for fp8 in (False, True):
        a, b = res[fp8]
        sig = np.sqrt(((a - b) ** 2).mean((0, 1)) / 2 / npx)
        assert np.all(np.abs(a.mean((0, 1)) - ref.mean((0, 1))) < 5 * sig + 1e-6), (fp8, a.mean((0, 1)), ref.mean((0, 1)))


def test_guided_graph_replay_gives_the_same_image(gpu_device):
    """Guided passes (inference, extended shading) replayed from a captured
    hipGraph follow the same random numbers as launched one by one: the same
    image up to FP64 summation order.  (With training the records arrive in
    atomic order, so trained networks differ between any two runs.)"""
    sc = scenes.guide_room(4, n_spheres=6, level=3, resolution=(48, 32))
    gs = wr.GpuScene(gpu_device, sc)
    print("guided render, graph vs launches: guide room 48x32", flush=True)
    g = wr.Guide(gs, seed=9)
    imgs = []
    for graph in (False, True):
        img, _, _ = _render(gs, 64, 12, seed=61, guide=g, guide_fraction=0.5, spp_per_pass=16, train_steps=0,
                            warmup_passes=0, use_graph=graph)
        imgs.append(img)
    import oracle
    print(f"  relMSE(graph, launches) {oracle.rel_mse(imgs[1], imgs[0]):.2e}")
    assert oracle.rel_mse(imgs[1], imgs[0]) < 1e-20
