"""Pins the FP64 guiding oracle (oracle/guide.py) by mathematics it cannot
satisfy by construction: finite differences, quadrature, closed forms and
hand-computed values.  CPU only."""
import numpy as np
import pytest

from oracle import guide as G


def random_records(rng, n, box=(-1, -1, -1, 1, 1, 1)):
    lo, hi = np.array(box[:3]), np.array(box[3:])
    pos = lo + (hi - lo) * rng.random((n, 3))

    def unit(k):
        v = rng.normal(size=(k, 3))
        return v / np.linalg.norm(v, axis=1, keepdims=True)

    return pos, unit(n), unit(n), unit(n), rng.random(n) + 0.1


def random_params(rng, scale_out=0.5):
    layers = []
    for l, (k, n) in enumerate(G.SHAPES):
        lim = np.sqrt(6.0 / k) * (scale_out if l == G.N_LAYERS - 1 else 1.0)
        layers.append((rng.uniform(-lim, lim, (k, n)), rng.uniform(-0.1, 0.1, n)))
    return G.pack(layers)


def test_parameter_layout():
    assert G.NUM_PARAMS == 17192
    p = np.arange(G.NUM_PARAMS, dtype=np.float64)
    layers = G.unpack(p)
    assert layers[0][0].shape == (32, 64) and layers[4][0].shape == (64, 40)
    assert layers[0][1][0] == 32 * 64 and layers[1][0][0, 0] == 32 * 64 + 64
    np.testing.assert_array_equal(G.pack(layers), p)


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, 1.0 + 2**-9, -2.5, 3.140625, 1e-3], dtype=np.float32)
    r = G.bf16(x)
    # 8 significant bits: ties go to even
    assert r[0] == 1.0
    assert r[1] == 1.0                  # tie between 1 and 1 + 2^-7 -> even (1)
    assert r[2] == 1.0 + 2 * 2**-7      # tie between 1 + 2^-7 and 1 + 2^-6 -> even
    assert r[3] == 1.0                  # below the tie
    assert r[4] == -2.5 and r[5] == 3.140625
    assert abs(r[6] - 1e-3) <= 1e-3 * 2**-8


def test_features_one_blob():
    # a value at bin centre j gives 1 there and exp(-1/2) at the neighbours
    enc = np.zeros((1, 7))
    enc[0, 0] = 2.5 / 6          # centre of bin 2 of 6
    enc[0, 3] = 0.5 / 4          # centre of bin 0 of 4
    enc[0, 6] = 1.0              # edge: distance 1/2 bin from the last centre of 3
    f = G.features(enc)[0]
    assert f.shape == (32,)
    np.testing.assert_allclose(f[0:6], np.exp(-0.5 * (np.arange(6) - 2.0) ** 2), rtol=1e-15)
    np.testing.assert_allclose(f[18:22], np.exp(-0.5 * np.arange(4.0) ** 2), rtol=1e-15)
    np.testing.assert_allclose(f[29:32], np.exp(-0.5 * (np.array([2.5, 1.5, 0.5])) ** 2), rtol=1e-15)


def fibonacci_sphere(n):
    i = np.arange(n) + 0.5
    y = 1 - 2 * i / n
    r = np.sqrt(1 - y * y)
    phi = np.pi * (3 - np.sqrt(5)) * i
    return np.column_stack([r * np.cos(phi), y, r * np.sin(phi)])


def raw_outputs(rng, lams):
    o = np.zeros((1, 40))
    for k in range(8):
        o[0, 5 * k] = rng.normal()
        o[0, 5 * k + 1:5 * k + 4] = rng.normal(size=3)
        o[0, 5 * k + 4] = np.log(np.expm1(lams[k]))  # softplus^-1
    return o


@pytest.mark.parametrize("lams", [[0.1] * 8, [1, 2, 3, 5, 8, 13, 21, 34], [50.0] * 8])
def test_pdf_integrates_to_one(lams):
    rng = np.random.default_rng(1)
    o = raw_outputs(rng, lams)
    d = fibonacci_sphere(400_000)
    p = G.pdf(np.repeat(o, len(d), axis=0), d)
    assert abs(p.mean() * 4 * np.pi - 1.0) < 2e-3


def test_single_lobe_closed_form_and_uniform_limit():
    # one dominant lobe: q = lam / (2 pi (1 - e^{-2 lam})) e^{lam (mu.w - 1)}
    o = np.zeros((1, 40))
    o[0, 0::5] = -200.0
    o[0, 0] = 0.0
    o[0, 1:4] = [0.0, 0.0, 2.0]
    lam = 7.0
    o[0, 4] = np.log(np.expm1(lam))
    w = np.array([[0.6, 0.0, 0.8]])
    expect = lam / (2 * np.pi * (1 - np.exp(-2 * lam))) * np.exp(lam * (0.8 - 1.0))
    np.testing.assert_allclose(G.pdf(o, w), expect, rtol=1e-12)
    # sharpness at the lower clamp: nearly uniform, 1/(4 pi) to first order in lambda
    o2 = np.zeros((1, 40))
    o2[0, 4::5] = -50.0  # softplus -> ~0, clamped to 0.05
    p = G.pdf(np.repeat(o2, 3, axis=0), np.eye(3))
    assert np.all(np.abs(p * 4 * np.pi - 1.0) < 0.06)


def test_sample_matches_pdf():
    rng = np.random.default_rng(2)
    o = raw_outputs(rng, [0.5, 2, 4, 8, 16, 1, 3, 30])
    n = 400_000
    d, _ = G.sample(np.repeat(o, n, axis=0), rng.random((n, 3)))
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)
    grid = fibonacci_sphere(400_000)
    q = G.pdf(np.repeat(o, len(grid), axis=0), grid)
    a = np.array([0.3, -0.5, 0.8])
    for f in [lambda v: v[:, 0], lambda v: v[:, 1] ** 2, lambda v: np.exp(v @ a)]:
        mc = f(d).mean()
        quad = (f(grid) * q).mean() * 4 * np.pi
        sd = f(d).std() / np.sqrt(n)
        assert abs(mc - quad) < 5 * sd + 1e-4, (mc, quad, sd)


def test_gradient_finite_differences():
    rng = np.random.default_rng(3)
    box = (-1, -1, -1, 1, 1, 1)
    pos, nrm, wo, wi, w = random_records(rng, 24, box)
    params = random_params(rng, scale_out=1.0)
    loss, grad = G.loss_and_grad(params, pos, nrm, wo, wi, w, box)
    assert np.isfinite(loss) and grad.shape == (G.NUM_PARAMS,)
    # every layer's weights and biases, and output-layer entries of every kind
    idx = []
    o = 0
    for k, n in G.SHAPES:
        idx += list(rng.choice(k * n, 6, replace=False) + o)
        idx += list(rng.choice(n, 3, replace=False) + o + k * n)
        o += k * n + n
    b4 = G.NUM_PARAMS - 40
    idx += [b4 + 5 * k + j for k in (0, 3, 7) for j in range(5)]
    h = 1e-6
    for i in idx:
        e = np.zeros_like(params)
        e[i] = h
        lp, _ = G.loss_and_grad(params + e, pos, nrm, wo, wi, w, box)
        lm, _ = G.loss_and_grad(params - e, pos, nrm, wo, wi, w, box)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (i, fd, grad[i])


def test_gradient_sign_descends():
    # a plausible sign error would make one small step increase the loss
    rng = np.random.default_rng(4)
    box = (-1, -1, -1, 1, 1, 1)
    pos, nrm, wo, wi, w = random_records(rng, 64, box)
    params = random_params(rng)
    l0, g = G.loss_and_grad(params, pos, nrm, wo, wi, w, box)
    l1, _ = G.loss_and_grad(params - 1e-3 * g / np.linalg.norm(g), pos, nrm, wo, wi, w, box)
    assert l1 < l0


def test_adam_scalar_by_hand():
    p, m, v = np.array([1.0]), np.array([0.0]), np.array([0.0])
    g = np.array([0.5])
    p1, m1, v1 = G.adam_step(p, m, v, g, 1, lr=0.1)
    # t = 1: m_hat = g, v_hat = g^2 -> step = lr * g / (|g| + eps)
    np.testing.assert_allclose(m1, 0.05)
    np.testing.assert_allclose(v1, 0.00025)
    np.testing.assert_allclose(p1, 1.0 - 0.1 * 0.5 / (0.5 + 1e-8), rtol=1e-15)
    p2, _, _ = G.adam_step(p1, m1, v1, np.array([-0.25]), 2, lr=0.1)
    m2 = 0.9 * 0.05 + 0.1 * -0.25
    v2 = 0.999 * 0.00025 + 0.001 * 0.0625
    step = 0.1 * (m2 / (1 - 0.81)) / (np.sqrt(v2 / (1 - 0.999**2)) + 1e-8)
    np.testing.assert_allclose(p2, p1 - step, rtol=1e-14)


def test_bf16_forward_close_to_fp64():
    # the BF16 emulation is a perturbation of the FP64 network, not another network
    rng = np.random.default_rng(5)
    box = (-1, -1, -1, 1, 1, 1)
    pos, nrm, wo, _, _ = random_records(rng, 128, box)
    params = random_params(rng)
    a, _ = G.forward(params, pos, nrm, wo, box)
    b, _ = G.forward(params, pos, nrm, wo, box, bf=True)
    rel = np.linalg.norm(a - b) / np.linalg.norm(a)
    assert 1e-5 < rel < 3e-2


def test_e4m3_known_values():
    q = G.e4m3
    assert q(1.0) == 1.0 and q(-2.5) == -2.5 and q(448.0) == 448.0 and q(0.0) == 0.0
    assert q(1.0625) == 1.0            # tie between 1 and 1.125 -> even mantissa (1.0)
    assert q(1.1875) == 1.25           # tie between 1.125 and 1.25 -> even (1.25)
    assert q(1.19) == 1.25 and q(1.18) == 1.125
    assert q(500.0) == 448.0 and q(-1e6) == -448.0   # saturation
    assert q(2.0 ** -9) == 2.0 ** -9   # smallest subnormal
    assert q(0.75 * 2.0 ** -9) == 2.0 ** -9 and q(0.5 * 2.0 ** -9) == 0.0  # tie to even (0)
    assert q(1.5 * 2.0 ** -9) == 2.0 * 2.0 ** -9     # tie between 1 and 2 steps -> 2
    assert q(8.5 * 2.0 ** -9) == 2.0 ** -6           # tie at the normal boundary -> 8 steps
    assert q(240.0) == 240.0 and q(232.0) == 224.0   # 232: tie between 224 and 240 -> 224 (even)
    # codes 0x01..0x7E (0x7F is NaN): 7 subnormals, 14 binades of 8, and 7 in [256, 448]
    grid = np.exp(np.linspace(np.log(2.0 ** -10), np.log(460.0), 400_000))
    vals = np.unique(q(grid))
    assert len(vals[vals > 0]) == 126 and vals.max() == 448.0


def test_pow2_exponent_range():
    assert G.pow2_exponent(1.0) == -7 and G.pow2_exponent(1.5) == -6 and G.pow2_exponent(128.0) == 0
    assert G.pow2_exponent(0.0) == 0 and G.pow2_exponent(129.0) == 1
    m = np.exp(np.random.default_rng(6).uniform(-40, 40, 10000)).astype(np.float32)
    k = G.pow2_exponent(m)
    s = m.astype(np.float64) * 2.0 ** -k.astype(np.float64)
    assert np.all((s > 64) & (s <= 128))


def test_fp8_forward_reduces_to_fp64_and_is_a_bounded_perturbation():
    rng = np.random.default_rng(7)
    box = (-1, -1, -1, 1, 1, 1)
    pos, nrm, wo, _, _ = random_records(rng, 200, box)
    params = random_params(rng).astype(np.float32)
    ref, _ = G.forward(params.astype(np.float64), pos, nrm, wo, box)
    wide = G.forward8(params, pos, nrm, wo, box, mbits=23)
    # 23-bit mantissa: only the FP32 rounding of features / hidden values remains
    assert np.abs(wide - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())
    out8 = G.forward8(params, pos, nrm, wo, box)
    bf, _ = G.forward(params.astype(np.float64), pos, nrm, wo, box, bf=True)
    e8 = np.linalg.norm(out8 - ref) / np.linalg.norm(ref)
    ebf = np.linalg.norm(bf - ref) / np.linalg.norm(ref)
    # 3 mantissa bits against BF16's 8: about 2^5 times the error, and still small
    assert 4 * ebf < e8 < 0.15, (e8, ebf)
    # a record's outputs do not depend on the other records
    np.testing.assert_array_equal(G.forward8(params, pos[:1], nrm[:1], wo[:1], box), out8[:1])
