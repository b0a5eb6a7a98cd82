"""Closed-form expectations shared by the oracle and GPU traversal tests."""
import numpy as np

from wurstrenderer_amd import scenes

def quad_scene_and_rays(n=6000, seed=7):
    """Four tilted quads (two triangles each) one behind the other, and rays
    from near the origin toward random points around them, with each ray's
    closest hit worked out in closed form: o + t d = c + a u + b v solved in
    FP64 per quad, a hit when |a| < 1 and |b| < 1 and t > 0, triangle 0 of the
    quad when a > b (scenes.quad's split), the smallest t winning.  Rays whose
    (a, b) on some quad lies within 1e-5 of an edge or of the diagonal are
    marked ambiguous."""
    rng = np.random.default_rng(seed)
    b = scenes.SceneBuilder()
    m = b.material(scenes.LAMBERT, rho_d=(0.5, 0.5, 0.5))
    quads = []
    for k in range(4):
        c = np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), 2.0 + 2.0 * k])
        tilt = rng.uniform(-0.5, 0.5, 2)
        u = np.array([np.cos(tilt[0]), 0.0, np.sin(tilt[0])]) * rng.uniform(0.4, 0.8)
        v = np.array([0.0, np.cos(tilt[1]), np.sin(tilt[1])]) * rng.uniform(0.4, 0.8)
        verts, faces = scenes.quad(c, u, v)
        verts = verts.astype(np.float32)
        b.mesh(verts, faces, material=m)
        v64 = verts.astype(np.float64)
        quads.append((v64[0] + 0.5 * (v64[2] - v64[0]), 0.5 * (v64[1] - v64[0]), 0.5 * (v64[3] - v64[0])))
    b.mesh(*scenes.quad((0, 0, -50), (1, 0, 0), (0, 1, 0)), light_L=(1, 1, 1))
    sc = b.build(dict(pos=(0, 0, -1), lookat=(0, 0, 1), fovy=40.0, resolution=(8, 8)), "quads")
    o = np.column_stack([rng.uniform(-0.5, 0.5, n), rng.uniform(-0.5, 0.5, n), rng.uniform(-0.5, 0.5, n)])
    target = np.column_stack([rng.uniform(-1.2, 1.2, n), rng.uniform(-1.2, 1.2, n), rng.uniform(1.5, 9.0, n)])
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = o
    rays[:, 3] = 0.0
    rays[:, 4:7] = d
    rays[:, 7] = np.inf
    o64, d64 = rays[:, :3].astype(np.float64), rays[:, 4:7].astype(np.float64)
    geom = np.full(n, 0xFFFFFFFF, np.uint32)
    prim = np.zeros(n, np.uint32)
    t = np.full(n, np.inf)
    amb = np.zeros(n, bool)
    for k, (c, u, v) in enumerate(quads):
        # [d, -u, -v] (t, a, b)^T = c - o
        M = np.stack([d64, np.broadcast_to(-u, d64.shape), np.broadcast_to(-v, d64.shape)], axis=2)
        tab = np.linalg.solve(M, (c - o64)[:, :, None])[:, :, 0]
        tk, a, bb = tab[:, 0], tab[:, 1], tab[:, 2]
        near = (np.abs(np.abs(a) - 1) < 1e-5) | (np.abs(np.abs(bb) - 1) < 1e-5) | (np.abs(a - bb) < 1e-5)
        amb |= near & (np.abs(a) < 1 + 1e-5) & (np.abs(bb) < 1 + 1e-5) & (tk > 0)
        hit = (np.abs(a) < 1) & (np.abs(bb) < 1) & (tk > 0) & (tk < t)
        geom[hit] = k
        prim[hit] = np.where(a[hit] > bb[hit], 0, 1)
        t[hit] = tk[hit]
    return sc, rays, geom, prim, t, amb

