"""FP64 restatement of the path-guiding network, its mixture and its training.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
network (wurstrenderer_amd/csrc/wr_guide.h, wr_guide.hip).  The product never
imports it.

No reference counterpart: the reference has no learned guiding.  Its only
spherical-Gaussian convention is SgLobe::eval = a exp((1 - w.mu) lambda)
(src/sphericalrepresentation/sphericalgaussian.h:10-21), which peaks at
w = -mu; this restatement uses the normalised lobe
    c(w) = lambda / (2 pi (1 - exp(-2 lambda))) exp(lambda (mu.w - 1)).

Pinning (tests/test_oracle_guide.py):
  - backward(): central finite differences of the FP64 loss (all layers,
    every output kind: logits, axes, sharpness);
  - pdf(): integrates to 1 over the sphere (quadrature), reduces to the
    closed-form vMF for one lobe, to 1/(4 pi) as lambda -> 0;
  - sample(): sample means of test functions match quadrature of pdf();
  - features(): one-blob values against the Gaussian-kernel definition at
    bin centres and edges;
  - adam_step(): against the update written out for a scalar by hand;
  - e4m3(): hand-computed values of the OCP E4M3 format (ties to even,
    subnormals, saturation, 126 positive finite values);
  - pow2_exponent(): the (64, 128] range it promises, on hand and random values;
  - forward8(): with the quantiser's mantissa widened to 23 bits it reduces to
    forward() (pins the per-layer and per-record scale bookkeeping), and at 3
    bits it is a bounded perturbation of it.
With bf16=True every rounding point of the GPU (weights, features, hidden
activations, output gradients) is rounded to BF16 (round to nearest even), so
the GPU's FP32-accumulated results agree to accumulation order.  forward8()
rounds at the GPU's FP8 points the same way.
"""
from __future__ import annotations

import numpy as np

K_IN, K_HID, K_OUT, N_LAYERS, N_LOBES = 32, 64, 40, 5, 8
LAMBDA_MIN, LAMBDA_MAX = 0.05, 2000.0
SHAPES = [(K_IN, K_HID), (K_HID, K_HID), (K_HID, K_HID), (K_HID, K_HID), (K_HID, K_OUT)]
NUM_PARAMS = sum(k * n + n for k, n in SHAPES)
assert NUM_PARAMS == 17192


def bf16(x):
    """FP32 -> BF16 (round to nearest even), returned as float64."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    return ((u + r) & 0xFFFF0000).astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    """Parameter vector -> [(W [K,N], b [N])] (W row-major, input-major)."""
    p = np.asarray(params, dtype=np.float64)
    out, o = [], 0
    for k, n in SHAPES:
        w = p[o:o + k * n].reshape(k, n)
        o += k * n
        b = p[o:o + n]
        o += n
        out.append((w, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([w.ravel(), b]) for w, b in layers])


# ------------------------------------------------------------------ encoding
def oct01(d):
    """Octahedral map of unit vectors to [0,1]^2 (Engelhardt and Dachsbacher 2008), y up:
    u = x / |d|_1, v = z / |d|_1, the lower half folded over the diagonals."""
    d = np.asarray(d, dtype=np.float64)
    l1 = np.abs(d).sum(axis=1)
    u, v = d[:, 0] / l1, d[:, 2] / l1
    sgn = lambda t: np.where(t >= 0, 1.0, -1.0)
    low = d[:, 1] < 0
    uu = np.where(low, (1 - np.abs(v)) * sgn(u), u)
    vv = np.where(low, (1 - np.abs(u)) * sgn(v), v)
    return np.clip(0.5 * uu + 0.5, 0, 1), np.clip(0.5 * vv + 0.5, 0, 1)


def encode(pos, nrm, wo, box):
    """(n,7) in [0,1]: position in the box (3), normal (2), direction (2)."""
    pos = np.asarray(pos, dtype=np.float64)
    lo, hi = np.asarray(box[:3], dtype=np.float64), np.asarray(box[3:], dtype=np.float64)
    p01 = np.clip((pos - lo) / np.maximum(hi - lo, 1e-20), 0.0, 1.0)
    na, nb = oct01(nrm)
    oa, ob = oct01(wo)
    return np.column_stack([p01, na, nb, oa, ob])


def features(enc):
    """One-blob encoding (Mueller et al. 2019): kernel exp(-(k (x - c_j))^2 / 2)
    over k bins with centres c_j = (j + 1/2) / k; 3 x 6 + 2 x 4 + 2 x 3 = 32."""
    cols = []
    for dim, k in [(0, 6), (1, 6), (2, 6), (3, 4), (4, 4), (5, 3), (6, 3)]:
        x = enc[:, dim:dim + 1]
        c = (np.arange(k) + 0.5) / k
        cols.append(np.exp(-0.5 * ((x - c) * k) ** 2))
    return np.concatenate(cols, axis=1)


# ------------------------------------------------------------------ network
def forward(params, pos, nrm, wo, box, bf=False):
    """Raw outputs (n,40) and the cache for backward()."""
    rnd = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(params)
    x = rnd(features(encode(pos, nrm, wo, box)).astype(np.float32) if bf else features(encode(pos, nrm, wo, box)))
    xs = [x]
    for l, (w, b) in enumerate(layers):
        y = x @ rnd(w) + b
        if l < N_LAYERS - 1:
            x = rnd(np.maximum(y, 0.0))
            xs.append(x)
        else:
            out = y
    return out, xs


def softplus(x):
    return np.where(x > 30, x, np.log1p(np.exp(np.minimum(x, 30))))


def log_vmf_norm(lam):
    return np.log(lam) - np.log(2 * np.pi) - np.log(-np.expm1(-2 * lam))


def decode(out):
    """(log w (n,8), mu (n,8,3), lambda (n,8)) of raw outputs (n,40)."""
    o = np.asarray(out, dtype=np.float64).reshape(-1, N_LOBES, 5)
    logits = o[:, :, 0]
    m = logits.max(axis=1, keepdims=True)
    lse = m + np.log(np.exp(logits - m).sum(axis=1, keepdims=True))
    a = o[:, :, 1:4]
    nrm = np.linalg.norm(a, axis=2, keepdims=True)
    mu = np.where(nrm > 1e-12, a / np.maximum(nrm, 1e-300), np.array([0.0, 1.0, 0.0]))
    lam = np.clip(softplus(o[:, :, 4]), LAMBDA_MIN, LAMBDA_MAX)
    return logits - lse, mu, lam


def log_pdf(out, dirs):
    logw, mu, lam = decode(out)
    d = np.asarray(dirs, dtype=np.float64)
    lc = logw + log_vmf_norm(lam) + lam * (np.einsum("nkj,nj->nk", mu, d) - 1.0)
    m = lc.max(axis=1, keepdims=True)
    return (m + np.log(np.exp(lc - m).sum(axis=1, keepdims=True)))[:, 0]


def pdf(out, dirs):
    return np.exp(log_pdf(out, dirs))


def frame(n):
    """Duff et al. 2017 ONB with y = n (as wr_device.h make_frame)."""
    sign = np.where(n[:, 1] >= 0, 1.0, -1.0)
    a = -1.0 / (sign + n[:, 1])
    b = n[:, 2] * n[:, 0] * a
    x = np.column_stack([sign + n[:, 0] ** 2 * a, -n[:, 0], b])
    z = np.column_stack([sign * b, -sign * n[:, 2], 1.0 + sign * n[:, 2] ** 2 * a])
    return x, n, z


def sample(out, u):
    """One direction per row from uniforms u (n,3): lobe by weight (u0), vMF around
    its axis (Wood 1994): cos = 1 + log(1 - u1 + u1 e^{-2 lambda}) / lambda, phi = 2 pi u2."""
    logw, mu, lam = decode(out)
    u = np.asarray(u, dtype=np.float64)
    c = np.cumsum(np.exp(logw), axis=1)
    k = np.minimum((u[:, :1] >= c[:, :-1]).sum(axis=1), N_LOBES - 1)
    rows = np.arange(len(u))
    m, l = mu[rows, k], lam[rows, k]
    ct = np.clip(1.0 + np.log(1.0 - u[:, 1] + u[:, 1] * np.exp(-2 * l)) / l, -1.0, 1.0)
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    phi = 2 * np.pi * u[:, 2]
    fx, fy, fz = frame(m)
    d = fx * (st * np.cos(phi))[:, None] + fy * ct[:, None] + fz * (st * np.sin(phi))[:, None]
    return d / np.linalg.norm(d, axis=1, keepdims=True), k


# ------------------------------------------------------------------ loss
def loss_grad_out(out, wi, w):
    """Per-record loss -w log q(wi) and its gradient w.r.t. the raw outputs (analytic)."""
    o = np.asarray(out, dtype=np.float64).reshape(-1, N_LOBES, 5)
    wi = np.asarray(wi, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    logw, mu, lam = decode(out)
    raw = o[:, :, 4]
    sp = softplus(raw)
    dl = np.where((sp > LAMBDA_MIN) & (sp < LAMBDA_MAX), 1.0 / (1.0 + np.exp(-raw)), 0.0)
    cs = np.einsum("nkj,nj->nk", mu, wi)
    lc = logw + log_vmf_norm(lam) + lam * (cs - 1.0)
    m = lc.max(axis=1, keepdims=True)
    logq = m + np.log(np.exp(lc - m).sum(axis=1, keepdims=True))
    gam = np.exp(lc - logq)
    g = np.zeros_like(o)
    g[:, :, 0] = gam - np.exp(logw)
    dlogc = 1.0 / lam - 2.0 / np.expm1(2 * lam) + (cs - 1.0)
    g[:, :, 4] = gam * dlogc * dl
    a = o[:, :, 1:4]
    nrm = np.linalg.norm(a, axis=2, keepdims=True)
    gm = wi[:, None, :] * (gam * lam)[:, :, None]
    ga = (gm - mu * np.einsum("nkj,nkj->nk", mu, gm)[:, :, None]) / np.maximum(nrm, 1e-300)
    g[:, :, 1:4] = np.where(nrm > 1e-12, ga, 0.0)
    g *= -w[:, None, None]
    return -w * logq[:, 0], g.reshape(-1, K_OUT)


def backward(params, xs, gout, bf=False):
    """Parameter gradient from the forward cache and d loss / d outputs."""
    rnd = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(params)
    grads = [None] * N_LAYERS
    dy = rnd(gout)
    for l in range(N_LAYERS - 1, -1, -1):
        w, _ = layers[l]
        x = xs[l]
        grads[l] = (x.T @ dy, dy.sum(axis=0))
        if l > 0:
            dx = dy @ rnd(w).T
            dy = rnd(dx * (x > 0))
    return pack(grads)


def loss_and_grad(params, pos, nrm, wo, wi, weight, box, scale=None, bf=False):
    """Mean loss (1/n) sum -weight log q(wi) and its parameter gradient."""
    n = len(pos)
    scale = 1.0 / n if scale is None else scale
    out, xs = forward(params, pos, nrm, wo, box, bf=bf)
    w = np.asarray(weight, dtype=np.float64) * scale
    if bf:  # the GPU scales the FP32 weight in FP32
        w = (np.asarray(weight, dtype=np.float32) * np.float32(scale)).astype(np.float64)
    loss, gout = loss_grad_out(out, wi, w)
    if bf:
        gout = gout.astype(np.float32)
    return loss.sum(), backward(params, xs, gout, bf=bf)


def adam_step(params, m, v, g, t, lr=1e-2, beta1=0.9, beta2=0.999, eps=1e-8):
    """Adam (Kingma and Ba 2015), bias-corrected; returns (params, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


# ------------------------------------------------------------------ FP8 (e4m3) inference
E4M3_MAX = 448.0


def e4m3(x, mbits=3):
    """Round to the OCP E4M3 grid (bias 7: normals down to 2^-6, subnormals in
    steps of 2^-9, largest finite 448), nearest even, saturating.  mbits widens
    the mantissa (same exponent range) for the pinning test."""
    x = np.asarray(x, dtype=np.float64)
    a = np.minimum(np.abs(x), E4M3_MAX if mbits == 3 else np.inf)
    _, E = np.frexp(a)                    # a = f 2^E, f in [0.5, 1)
    e = np.maximum(E - 1, -6)             # binade; below 2^-6 the subnormal spacing
    step = np.ldexp(1.0, e - mbits)
    q = np.round(a / step) * step         # np.round: half to even
    if mbits == 3:
        q = np.minimum(q, E4M3_MAX)
    return np.where(a == 0, 0.0, np.copysign(q, x))


def pow2_exponent(m):
    """k = ceil(log2 m) - 7 from the FP32 bits of m (so m 2^-k is in (64, 128]),
    0 for m <= 0, clamped to [-100, 100]."""
    m = np.asarray(m, dtype=np.float32)
    u = m.view(np.uint32).astype(np.int64)
    e = ((u >> 23) & 0xFF) - 127 + ((u & 0x7FFFFF) != 0)
    k = np.clip(e - 7, -100, 100)
    return np.where(m > 0, k, 0)


def forward8(params, pos, nrm, wo, box, mbits=3):
    """Raw outputs (n, 40) of the FP8 inference variant: e4m3 weights with a
    power-of-two scale per layer (from the layer's largest |weight|), e4m3
    inputs with a power-of-two scale per layer and record (from the record's
    largest input to that layer), exact accumulation, FP32 biases."""
    layers = unpack(np.asarray(params, dtype=np.float32).astype(np.float64))
    x = features(encode(pos, nrm, wo, box)).astype(np.float32).astype(np.float64)
    for l, (w, b) in enumerate(layers):
        kw = int(pow2_exponent(np.float32(np.abs(w).max())))
        w8 = e4m3(np.ldexp(w, -kw), mbits)
        kx = pow2_exponent(x.astype(np.float32).max(axis=1))       # per record
        x8 = e4m3(np.ldexp(x, -kx[:, None]), mbits)
        y = (x8 @ w8) * np.ldexp(1.0, kw + kx)[:, None] + b
        if l < N_LAYERS - 1:
            x = np.maximum(y, 0.0).astype(np.float32).astype(np.float64)
        else:
            return y

