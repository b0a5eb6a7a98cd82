"""Seeded synthetic scenes for the `nextevent_path_tracer` workloads.

The reference loads OBJ meshes through Assimp (src/sceneio/simple2parser.h:27-68,
src/sceneio/meshio/assimpio.h:61-153); here scenes are generated from a seed so
the same triangles reach the GPU path, the oracle and the benchmark.  One
geometry = one mesh of the reference (its geomID is the attach order,
src/accel/embreeaccel.h:312); triangles of a geometry are contiguous.

Materials follow the reference's OBJ mapping (assimpio.h:79-121): Kd only ->
TwoSided(Lambert), Ks only -> TwoSided(Phong), both -> TwoSided(Mixed).  Emitters
are DiffuseAreaLight meshes without a BSDF (assimpio.h:128-139).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

LAMBERT, OREN_NAYAR, PHONG, MIXED = 0, 1, 2, 3

MATERIAL_DTYPE = np.dtype([("type", "<i4"), ("twosided", "<i4"), ("rho_d", "<f4", 3),
                           ("rho_s", "<f4", 3), ("sigma", "<f4"), ("exponent", "<f4")])
LIGHT_DTYPE = np.dtype([("geom", "<i4"), ("L", "<f4", 3)])
CAMERA_DTYPE = np.dtype([("pos", "<f4", 3), ("lookat", "<f4", 3), ("up", "<f4", 3), ("fovy_deg", "<f4"),
                         ("width", "<i4"), ("height", "<i4"), ("aperture_radius", "<f4"),
                         ("focal_distance", "<f4")])
assert MATERIAL_DTYPE.itemsize == 40 and LIGHT_DTYPE.itemsize == 16 and CAMERA_DTYPE.itemsize == 56


@dataclasses.dataclass
class Scene:
    verts: np.ndarray          # (nv, 3) float32
    tris: np.ndarray           # (nt, 3) uint32
    tri_geom: np.ndarray       # (nt,) uint32
    geom_material: np.ndarray  # (ng,) int32, -1 for emitters
    geom_light: np.ndarray     # (ng,) int32, -1 if not emissive
    materials: np.ndarray      # MATERIAL_DTYPE
    lights: np.ndarray         # LIGHT_DTYPE
    camera: np.ndarray         # CAMERA_DTYPE, shape ()
    name: str = "scene"
    # homogeneous medium filling the scene: {"sigma_a": rgb, "sigma_s": rgb, "g": HG asymmetry} or None
    medium: dict | None = None
    # envmap light (EnvmapLight over an equirectangular image): float32 (H, 2H, 3) or None
    envmap: np.ndarray | None = None

    @property
    def n_tris(self) -> int:
        return int(self.tris.shape[0])

    @property
    def resolution(self) -> tuple[int, int]:
        return int(self.camera["width"]), int(self.camera["height"])

    def with_medium(self, sigma_a=(0.05, 0.05, 0.05), sigma_s=(0.4, 0.4, 0.4), g=0.7) -> "Scene":
        """The scene filled with a homogeneous medium (a number broadcasts to RGB)."""
        rgb = lambda v: tuple(float(x) for x in np.broadcast_to(np.asarray(v, np.float64), (3,)))
        return dataclasses.replace(self, medium={"sigma_a": rgb(sigma_a), "sigma_s": rgb(sigma_s), "g": float(g)})

    def with_envmap(self, rgb) -> "Scene":
        """The scene lit by an equirectangular envmap: an (H, 2H, 3) image, or an RGB
        triple for a constant sky (a 2 x 1 image)."""
        a = np.asarray(rgb, np.float32)
        if a.ndim == 1:
            a = np.broadcast_to(a, (1, 2, 3)).copy()
        return dataclasses.replace(self, envmap=np.ascontiguousarray(a))

    def with_camera(self, **kw) -> "Scene":
        cam = self.camera.copy()
        for k, v in kw.items():
            cam[k] = v
        return dataclasses.replace(self, camera=cam)


class SceneBuilder:
    """Accumulates geometries (meshes) in attach order."""

    def __init__(self):
        self._verts, self._tris, self._geom = [], [], []
        self._nv = 0
        self.geom_material: list[int] = []
        self.geom_light: list[int] = []
        self.materials: list[tuple] = []
        self.lights: list[tuple] = []

    def material(self, type_, rho_d=(0, 0, 0), rho_s=(0, 0, 0), sigma=0.0, exponent=0.0, twosided=True) -> int:
        self.materials.append((type_, int(twosided), tuple(rho_d), tuple(rho_s), sigma, exponent))
        return len(self.materials) - 1

    def mesh(self, verts, tris, material: int | None = None, light_L=None) -> int:
        verts = np.asarray(verts, np.float64).reshape(-1, 3)
        tris = np.asarray(tris, np.int64).reshape(-1, 3)
        g = len(self.geom_material)
        self._verts.append(verts)
        self._tris.append(tris + self._nv)
        self._geom.append(np.full(len(tris), g, np.uint32))
        self._nv += len(verts)
        if light_L is not None:
            self.lights.append((g, tuple(float(x) for x in light_L)))
            self.geom_material.append(-1)
            self.geom_light.append(len(self.lights) - 1)
        else:
            assert material is not None
            self.geom_material.append(material)
            self.geom_light.append(-1)
        return g

    def build(self, camera: dict, name: str) -> Scene:
        cam = np.zeros((), CAMERA_DTYPE)
        cam["pos"], cam["lookat"], cam["up"] = camera["pos"], camera["lookat"], camera.get("up", (0, 1, 0))
        cam["fovy_deg"] = camera.get("fovy", 90.0)
        cam["width"], cam["height"] = camera["resolution"]
        cam["aperture_radius"] = camera.get("aperture", 0.0)
        cam["focal_distance"] = camera.get("focal", 1.0)
        mats = np.zeros(len(self.materials), MATERIAL_DTYPE)
        for i, m in enumerate(self.materials):
            mats[i] = m
        lights = np.zeros(len(self.lights), LIGHT_DTYPE)
        for i, l in enumerate(self.lights):
            lights[i] = l
        return Scene(
            verts=np.ascontiguousarray(np.concatenate(self._verts).astype(np.float32)),
            tris=np.ascontiguousarray(np.concatenate(self._tris).astype(np.uint32)),
            tri_geom=np.ascontiguousarray(np.concatenate(self._geom)),
            geom_material=np.asarray(self.geom_material, np.int32),
            geom_light=np.asarray(self.geom_light, np.int32),
            materials=mats, lights=lights, camera=cam, name=name)


def quad(c, u, v):
    """Quad centred at c spanned by half-axes u, v; Ng = u x v."""
    c, u, v = (np.asarray(x, np.float64) for x in (c, u, v))
    verts = [c - u - v, c + u - v, c + u + v, c - u + v]
    return np.array(verts), np.array([[0, 1, 2], [0, 2, 3]])


def box(center, size, yaw):
    """Axis box rotated by yaw about +y: 8 vertices, 12 outward-wound triangles."""
    hx, hy, hz = (0.5 * s for s in size)
    corners = np.array([[x, y, z] for x in (-hx, hx) for y in (-hy, hy) for z in (-hz, hz)])
    c, s = math.cos(yaw), math.sin(yaw)
    rot = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    verts = corners @ rot.T + np.asarray(center)
    # index = 4*ix + 2*iy + iz
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for a, b, cc, d in faces:
        tris += [[a, b, cc], [a, cc, d]]
    return verts, np.array(tris)


_ICO_CACHE: dict[int, tuple[np.ndarray, np.ndarray]] = {}


def icosphere(level: int):
    """Unit icosphere: 20*4^level triangles, 10*4^level+2 vertices."""
    if level in _ICO_CACHE:
        return _ICO_CACHE[level]
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = np.array(f, np.int64)
    for _ in range(level):
        cache: dict[tuple[int, int], int] = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces.tolist():
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = np.array(nf, np.int64)
    out = (np.array(verts), faces)
    _ICO_CACHE[level] = out
    return out


def cornell(seed: int = 1, resolution=(256, 256)) -> Scene:
    """Cornell-style box: 5 quads of [-1,1]^3 (front open), two seeded boxes,
    0.5 x 0.5 ceiling light of radiance 17 facing down.  36 triangles."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    gray = b.material(LAMBERT, rho_d=(0.75, 0.75, 0.75))
    red = b.material(LAMBERT, rho_d=(0.63, 0.065, 0.05))
    green = b.material(LAMBERT, rho_d=(0.14, 0.45, 0.09))
    b.mesh(*quad((0, -1, 0), (0, 0, 1), (1, 0, 0)), material=gray)     # floor, Ng +y
    b.mesh(*quad((0, 1, 0), (1, 0, 0), (0, 0, 1)), material=gray)      # ceiling, Ng -y
    b.mesh(*quad((0, 0, -1), (1, 0, 0), (0, 1, 0)), material=gray)     # back, Ng +z
    b.mesh(*quad((-1, 0, 0), (0, 1, 0), (0, 0, 1)), material=red)      # left, Ng +x
    b.mesh(*quad((1, 0, 0), (0, 0, 1), (0, 1, 0)), material=green)     # right, Ng -x
    yaw = rng.uniform(0, math.pi / 2, size=2)
    b.mesh(*box((-0.35, -0.4, -0.3), (0.6, 1.2, 0.6), yaw[0]), material=gray)
    b.mesh(*box((0.4, -0.7, 0.35), (0.6, 0.6, 0.6), yaw[1]), material=gray)
    b.mesh(*quad((0, 0.999, 0), (0.25, 0, 0), (0, 0, 0.25)), light_L=(17, 17, 17))  # Ng = -y
    return b.build(dict(pos=(0, 0, 3.75), lookat=(0, 0, 0), up=(0, 1, 0), fovy=40.0, resolution=resolution),
                   name=f"cornell_s{seed}")


def icospheres(seed: int = 2, n_spheres: int = 64, level: int = 5, resolution=(1920, 1080)) -> Scene:
    """64 level-5 icospheres (1.31M triangles) over a 200 x 200 floor, two 1 x 1
    downward emitters of radiance 20; materials 50/50 Lambert / modified Phong."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    sv, sf = icosphere(level)
    floor_mat = b.material(LAMBERT, rho_d=(0.5, 0.5, 0.5))
    b.mesh(*quad((0, -6.5, 0), (0, 0, 100), (100, 0, 0)), material=floor_mat)
    centres = rng.uniform(-5, 5, size=(n_spheres, 3))
    radii = rng.uniform(0.3, 1.2, size=n_spheres)
    kinds = rng.random(n_spheres) < 0.5
    albedo = rng.uniform(0.1, 0.9, size=(n_spheres, 3))
    expo = np.exp(rng.uniform(math.log(10), math.log(1000), size=n_spheres))
    for i in range(n_spheres):
        if kinds[i]:
            m = b.material(LAMBERT, rho_d=tuple(albedo[i]))
        else:
            m = b.material(PHONG, rho_s=(0.8, 0.8, 0.8), exponent=float(expo[i]))
        b.mesh(sv * radii[i] + centres[i], sf, material=m)
    for x in (-3.0, 3.0):
        b.mesh(*quad((x, 8, 0), (0.5, 0, 0), (0, 0, 0.5)), light_L=(20, 20, 20))
    return b.build(dict(pos=(0, 2, 18), lookat=(0, 0, 0), up=(0, 1, 0), fovy=40.0, resolution=resolution),
                   name=f"icospheres_s{seed}_n{n_spheres}_l{level}")


def guide_room(seed: int = 4, n_spheres: int = 49, level: int = 5, resolution=(1920, 1080), gap: float = 0.2,
               L: float = 500.0) -> Scene:
    """A closed 4 x 2.5 x 4 room lit only through a gap x gap hole in the ceiling
    by a hidden gap-sized quad light of radiance L inside a closed attic above
    it; the room is filled with seeded icosphere clutter (49 level-5 spheres:
    1,003,520 triangles).  Almost every surface sees the light only indirectly,
    through the bright patch the gap throws on the floor: the hard case that
    guiding addresses.  Walls and clutter are two-sided; 70% of the clutter is
    Lambert, 30% modified Phong."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    grey = b.material(LAMBERT, rho_d=(0.7, 0.7, 0.7))
    red = b.material(LAMBERT, rho_d=(0.63, 0.065, 0.05))
    green = b.material(LAMBERT, rho_d=(0.14, 0.45, 0.09))
    X, Y = 2.0, 2.5
    b.mesh(*quad((0, 0, 0), (0, 0, X), (X, 0, 0)), material=grey)            # floor
    b.mesh(*quad((0, Y / 2, -X), (X, 0, 0), (0, Y / 2, 0)), material=grey)   # back
    b.mesh(*quad((0, Y / 2, X), (X, 0, 0), (0, Y / 2, 0)), material=grey)    # front (behind the camera)
    b.mesh(*quad((-X, Y / 2, 0), (0, 0, X), (0, Y / 2, 0)), material=red)    # left
    b.mesh(*quad((X, Y / 2, 0), (0, 0, X), (0, Y / 2, 0)), material=green)   # right
    h = 0.5 * gap
    w = 0.5 * (X - h)
    # ceiling around the hole [-h, h]^2
    b.mesh(*quad((0, Y, h + w), (X, 0, 0), (0, 0, w)), material=grey)
    b.mesh(*quad((0, Y, -h - w), (X, 0, 0), (0, 0, w)), material=grey)
    b.mesh(*quad((-h - w, Y, 0), (w, 0, 0), (0, 0, h)), material=grey)
    b.mesh(*quad((h + w, Y, 0), (w, 0, 0), (0, 0, h)), material=grey)
    # closed attic above the hole, light near its top facing down
    A, top = 3 * h, Y + 0.4
    mid = 0.5 * (Y + top)
    b.mesh(*quad((0, top, 0), (A, 0, 0), (0, 0, A)), material=grey)
    for sx in (-A, A):
        b.mesh(*quad((sx, mid, 0), (0, 0, A), (0, 0.2, 0)), material=grey)
    for sz in (-A, A):
        b.mesh(*quad((0, mid, sz), (A, 0, 0), (0, 0.2, 0)), material=grey)
    sv, sf = icosphere(level)
    centres = np.column_stack([rng.uniform(-1.7, 1.7, n_spheres), rng.uniform(0.3, 2.0, n_spheres),
                               rng.uniform(-1.7, 1.7, n_spheres)])
    radii = rng.uniform(0.12, 0.3, n_spheres)
    lambert = rng.random(n_spheres) < 0.7
    albedo = rng.uniform(0.2, 0.8, size=(n_spheres, 3))
    expo = np.exp(rng.uniform(math.log(10), math.log(200), size=n_spheres))
    for i in range(n_spheres):
        if lambert[i]:
            m = b.material(LAMBERT, rho_d=tuple(albedo[i]))
        else:
            m = b.material(PHONG, rho_s=(0.6, 0.6, 0.6), exponent=float(expo[i]))
        b.mesh(sv * radii[i] + centres[i], sf, material=m)
    b.mesh(*quad((0, top - 0.02, 0), (h, 0, 0), (0, 0, h)), light_L=(L, L, L))
    return b.build(dict(pos=(0, 1.3, 1.95), lookat=(0, 0.9, -1.0), up=(0, 1, 0), fovy=65.0, resolution=resolution),
                   name=f"guide_room_s{seed}_n{n_spheres}_l{level}")


def cornell_clutter(seed: int, n_tris: int, resolution=(256, 256)) -> Scene:
    """The Cornell-style box of `cornell` filled with seeded icosphere clutter
    of about n_tris triangles (the batched-workload generator): level-5, 4 or 3
    spheres depending on the budget, Lambert with random albedo."""
    base = cornell(seed, resolution)
    rng = np.random.default_rng(seed)
    level = 5 if n_tris >= 200_000 else (4 if n_tris >= 40_000 else 3)
    sv, sf = icosphere(level)
    per = len(sf)
    count = max(1, int(round(n_tris / per)))
    b = SceneBuilder()
    # rebuild: the box first (its geometries and materials), then the clutter
    for g in range(len(base.geom_material)):
        sel = base.tri_geom == g
        tris = base.tris[sel]
        used, inv = np.unique(tris.ravel(), return_inverse=True)
        verts = base.verts[used]
        if base.geom_light[g] >= 0:
            b.mesh(verts, inv.reshape(-1, 3), light_L=tuple(base.lights[base.geom_light[g]]["L"]))
        else:
            m = base.materials[base.geom_material[g]]
            b.mesh(verts, inv.reshape(-1, 3), material=b.material(int(m["type"]), rho_d=tuple(m["rho_d"]),
                                                                   rho_s=tuple(m["rho_s"]), sigma=float(m["sigma"]),
                                                                   exponent=float(m["exponent"]),
                                                                   twosided=bool(m["twosided"])))
    r = 0.25 * (1.0 / count) ** (1.0 / 3.0) + 0.02
    centres = rng.uniform(-0.85, 0.85, size=(count, 3))
    radii = rng.uniform(0.4 * r, r, size=count)
    albedo = rng.uniform(0.2, 0.8, size=(count, 3))
    for i in range(count):
        b.mesh(sv * radii[i] + centres[i], sf, material=b.material(LAMBERT, rho_d=tuple(albedo[i])))
    cam = base.camera
    return b.build(dict(pos=tuple(cam["pos"]), lookat=tuple(cam["lookat"]), up=tuple(cam["up"]),
                        fovy=float(cam["fovy_deg"]), resolution=resolution),
                   name=f"cornell_clutter_s{seed}_t{n_tris}")


def batch_workload(seed: int = 3, n_scenes: int = 32, tri_range=(10_000, 2_000_000),
                   resolutions=(128, 256, 512, 1024), spp_range=(16, 256), depth_range=(8, 16)):
    """The batched workload: scene i uses seed 1000 + i; triangle counts
    log-uniform in tri_range, square resolutions from `resolutions`, spp and
    max depth uniform integers.  Returns [(scene, spp, num_path_vertices)]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_scenes):
        t = int(math.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1]))))
        res = int(rng.choice(resolutions))
        spp = int(rng.integers(spp_range[0], spp_range[1] + 1))
        depth = int(rng.integers(depth_range[0], depth_range[1] + 1))
        out.append((cornell_clutter(1000 + i, t, (res, res)), spp, depth + 2))
    return out


def medium_clutter(seed: int = 5, resolution=(1280, 720), tri_range=(10_000, 2_000_000), sigma_a=0.05, sigma_s=0.4,
                   g=0.7) -> Scene:
    """The batched-workload scene type filled with a homogeneous medium: a
    triangle budget drawn log-uniform in tri_range from `seed`, the Cornell box
    with that much icosphere clutter (cornell_clutter with the same seed), and
    the medium (sigma_a, sigma_s RGB-equal, Henyey-Greenstein g) filling space."""
    rng = np.random.default_rng(seed)
    t = int(math.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1]))))
    sc = cornell_clutter(seed, t, resolution)
    return sc.with_medium(sigma_a=sigma_a, sigma_s=sigma_s, g=g)


def emitting_box(resolution=(24, 24), L=1.0, half=1.0) -> Scene:
    """Camera at the centre of a closed cube whose six faces emit L inward (no
    reflecting surface): with a purely absorbing medium a pixel sees
    L exp(-sigma_t d), d the distance to the wall; with a purely scattering one
    the radiance is L everywhere (radiative equilibrium)."""
    b = SceneBuilder()
    v, t = box((0, 0, 0), (2 * half,) * 3, 0.0)
    b.mesh(v, t[:, ::-1].copy(), light_L=(L, L, L))  # inward-wound: Ng points into the cube
    return b.build(dict(pos=(0.0, 0.0, 0.0), lookat=(0.0, 0.0, -1.0), up=(0, 1, 0), fovy=70.0,
                        resolution=resolution), name="emitting_box")


def box_exit_distance(o, d, half=1.0):
    """Distance from points o (inside) along unit directions d to the cube [-half, half]^3."""
    t = np.full(len(d), np.inf)
    for k in range(3):
        with np.errstate(divide="ignore", invalid="ignore"):
            tk = np.where(d[:, k] > 0, (half - o[:, k]) / d[:, k], (-half - o[:, k]) / d[:, k])
        t = np.minimum(t, np.where(d[:, k] != 0, tk, np.inf))
    return t


def plane_under_light(resolution=(32, 32), light_half=0.5, height=1.0, rho=0.8, L=10.0, plane_half=4.0) -> Scene:
    """Lambert plane y=0 under a square emitter at y=height facing down; the
    camera looks straight down from below the emitter.  Direct lighting of
    this scene has a closed form (Lambert's polygon formula)."""
    b = SceneBuilder()
    m = b.material(LAMBERT, rho_d=(rho, rho, rho))
    b.mesh(*quad((0, 0, 0), (0, 0, plane_half), (plane_half, 0, 0)), material=m)
    b.mesh(*quad((0, height, 0), (light_half, 0, 0), (0, 0, light_half)), light_L=(L, L, L))
    return b.build(dict(pos=(0, 0.6 * height, 0.0001), lookat=(0, 0, 0), up=(0, 0, -1), fovy=60.0,
                        resolution=resolution), name="plane_under_light")


def triangle_soup(n: int, seed: int = 0, extent: float = 4.0, size: float = 0.6) -> Scene:
    """Random triangles (several geometries) for traversal parity tests."""
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    m = b.material(LAMBERT, rho_d=(0.5, 0.5, 0.5))
    per = max(1, n // 4)
    left = n
    while left > 0:
        k = min(per, left)
        c = rng.uniform(-extent, extent, size=(k, 1, 3))
        v = c + rng.normal(scale=size, size=(k, 3, 3))
        b.mesh(v.reshape(-1, 3), np.arange(3 * k).reshape(-1, 3), material=m)
        left -= k
    b.mesh(*quad((0, extent + 2, 0), (0.5, 0, 0), (0, 0, 0.5)), light_L=(1, 1, 1))
    return b.build(dict(pos=(0, 0, 3 * extent), lookat=(0, 0, 0), fovy=45.0, resolution=(64, 64)),
                   name=f"soup_{n}_s{seed}")


def camera_rays(scene: Scene, spp: int = 1, seed: int = 0) -> np.ndarray:
    """Pinhole primary rays (8 floats: o, tmin, d, tmax) through jittered pixels."""
    cam = scene.camera
    W, H = scene.resolution
    pos = cam["pos"].astype(np.float64)
    z = cam["lookat"].astype(np.float64) - pos
    z /= np.linalg.norm(z)
    x = np.cross(cam["up"].astype(np.float64), z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    d = H * 0.5 / math.tan(math.radians(float(cam["fovy_deg"])) * 0.5)
    rng = np.random.default_rng(seed)
    py, px = np.mgrid[0:H, 0:W]
    px = np.repeat(px.reshape(-1), spp) + rng.random(W * H * spp)
    py = np.repeat(py.reshape(-1), spp) + rng.random(W * H * spp)
    dl = np.stack([-(px - 0.5 * W), -(py - 0.5 * H), np.full_like(px, d)], 1)
    dl /= np.linalg.norm(dl, axis=1, keepdims=True)
    dw = dl[:, :1] * x + dl[:, 1:2] * y + dl[:, 2:3] * z
    rays = np.zeros((len(dw), 8), np.float32)
    rays[:, 0:3] = pos
    rays[:, 3] = 1e-4
    rays[:, 4:7] = dw
    rays[:, 7] = np.inf
    return rays


def random_rays(scene: Scene, n: int, seed: int = 0, inside: bool = False) -> np.ndarray:
    """Rays from a sphere around the scene toward random interior points
    (inside=False), or from random points inside the bounds (inside=True)."""
    rng = np.random.default_rng(seed)
    lo, hi = scene.verts.min(0).astype(np.float64), scene.verts.max(0).astype(np.float64)
    c, r = 0.5 * (lo + hi), 0.5 * np.linalg.norm(hi - lo) + 1e-3
    if inside:
        o = rng.uniform(lo, hi, size=(n, 3))
        d = rng.normal(size=(n, 3))
    else:
        u = rng.normal(size=(n, 3))
        o = c + 1.5 * r * u / np.linalg.norm(u, axis=1, keepdims=True)
        tgt = rng.uniform(lo, hi, size=(n, 3))
        d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 1e-4
    rays[:, 4:7] = d
    rays[:, 7] = np.inf
    return rays


def write_obj(scene: Scene, directory: str, name: str = "scene") -> str:
    """Writes the scene as OBJ + MTL in the form the reference's loader maps
    (one object and material per geometry; Kd -> Lambert, Ks/Ns -> Phong with
    exponent Ns/4, Ke -> emitter) and returns the OBJ path.  Floats are written
    with 9 significant digits, so FP32 values round-trip exactly."""
    import os
    os.makedirs(directory, exist_ok=True)
    obj_path = os.path.join(directory, name + ".obj")
    mtl_lines, obj_lines = [], [f"mtllib {name}.mtl"]
    f9 = lambda x: format(float(x), ".9g")
    for m_i, m in enumerate(scene.materials):
        t = int(m["type"])
        if t == OREN_NAYAR or not int(m["twosided"]):
            raise ValueError("OBJ/MTL has no encoding for Oren-Nayar or one-sided materials")
        kd = m["rho_d"] if t in (LAMBERT, MIXED) else np.zeros(3)
        ks = m["rho_s"] if t in (PHONG, MIXED) else np.zeros(3)
        ns = 4.0 * float(m["exponent"]) if t in (PHONG, MIXED) else 0.0
        mtl_lines += [f"newmtl m{m_i}", "Kd " + " ".join(f9(x) for x in kd), "Ks " + " ".join(f9(x) for x in ks),
                      f"Ns {f9(ns)}", "Ke 0 0 0"]
    for l_i, l in enumerate(scene.lights):
        mtl_lines += [f"newmtl e{l_i}", "Kd 0 0 0", "Ks 0 0 0", "Ns 0", "Ke " + " ".join(f9(x) for x in l["L"])]
    for v in scene.verts:
        obj_lines.append("v " + " ".join(f9(x) for x in v))
    for g in range(len(scene.geom_material)):
        mat = f"m{scene.geom_material[g]}" if scene.geom_light[g] < 0 else f"e{scene.geom_light[g]}"
        obj_lines += [f"o geom{g}", f"usemtl {mat}"]
        for t in scene.tris[scene.tri_geom == g]:
            obj_lines.append("f " + " ".join(str(int(i) + 1) for i in t))
    with open(os.path.join(directory, name + ".mtl"), "w") as f:
        f.write("\n".join(mtl_lines) + "\n")
    with open(obj_path, "w") as f:
        f.write("\n".join(obj_lines) + "\n")
    if scene.envmap is not None:
        write_pfm(os.path.join(directory, f"{name}_env.pfm"), scene.envmap)
    return obj_path


def write_pfm(path: str, rgb) -> None:
    """A PFM image ("PF", little endian, rows bottom-up: floatimage.h:435-462) of
    an (H, W, 3) array whose row 0 is the top."""
    a = np.asarray(rgb, np.float32)
    with open(path, "wb") as f:
        f.write(f"PF\n{a.shape[1]} {a.shape[0]}\n-1\n".encode())
        f.write(np.ascontiguousarray(a[::-1]).astype("<f4").tobytes())


def scene_json(scene: Scene, obj_name: str, renders: list[dict]) -> dict:
    """The Simple2 scene JSON (src/sceneio/simple2parser.h:102-223 keys) for a
    scene written by write_obj."""
    cam = scene.camera
    j = {"camera": {"pos": [float(x) for x in cam["pos"]], "lookat": [float(x) for x in cam["lookat"]],
                    "up": [float(x) for x in cam["up"]], "resolution": [int(cam["width"]), int(cam["height"])],
                    "fovy": float(cam["fovy_deg"])},
         "meshes": [{"path": obj_name}], "render": renders}
    if float(cam["aperture_radius"]) > 0:
        if float(cam["aperture_radius"]) != float(cam["focal_distance"]):
            raise ValueError("the JSON key \"radius\" sets both the aperture and the focal distance")
        j["camera"]["radius"] = float(cam["aperture_radius"])
    if scene.envmap is not None:  # written by write_obj next to the OBJ
        j["envmap"] = obj_name[:-4] + "_env.pfm"
    if scene.medium:
        j["medium"] = {"type": "homogeneous", "sigma_a": list(scene.medium["sigma_a"]),
                       "sigma_s": list(scene.medium["sigma_s"]), "g": scene.medium["g"]}
    return j
