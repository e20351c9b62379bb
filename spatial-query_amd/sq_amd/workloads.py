"""Seeded synthetic scenes and query batches (host side, f32 arrays).

Every stream is SplitMix64 exactly as the reference tests draw it
(crates/spatial-query/tests/conformance.rs:9-25): draw k of a stream seeded
with `s` is mix(s + (k+1) * 0x9E3779B97F4A7C15), and `range(lo, hi)` is
`lo + ((x >> 11) as f32 / 2^53 as f32) * (hi - lo)` in f32 with no FMA.  Since
draw k depends only on k, whole streams are generated vectorised.

Unit directions are drawn by rejection from the cube [-1, 1]^D and normalised
with `Ray::new` (point.rs:82-89), so they use only f32 arithmetic and a
correctly rounded sqrt: the arrays are bit-identical on every machine.  The
heightfield and the camera use sin/cos (float64 numpy, rounded to f32 once);
the engine and the checker always consume the same arrays.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
TWO53 = np.float32(2.0 ** 53)


def splitmix(seed: int, n: int, start: int = 0) -> np.ndarray:
    """Draws start .. start+n-1 of the SplitMix64 stream seeded with `seed`."""
    k = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    return z


def u01(x: np.ndarray) -> np.ndarray:
    """(x >> 11) as f32 / 2^53 as f32 (round-to-nearest u64 -> f32, then exact scaling)."""
    return (x >> np.uint64(11)).astype(np.float32) / TWO53


def frange(x: np.ndarray, lo: float, hi: float) -> np.ndarray:
    lo32, hi32 = np.float32(lo), np.float32(hi)
    return lo32 + u01(x) * (hi32 - lo32)


class Stream:
    """Sequential draws from one SplitMix64 stream (vectorised in blocks)."""

    def __init__(self, seed: int):
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.pos = 0

    def u64(self, n: int) -> np.ndarray:
        out = splitmix(self.seed, n, self.pos)
        self.pos += n
        return out

    def range(self, lo: float, hi: float, n: int) -> np.ndarray:
        return frange(self.u64(n), lo, hi)


def normalize_or_zero(v: np.ndarray) -> np.ndarray:
    """Point::normalize_or_zero in f32 (point.rs:60-89): ((0+x*x)+y*y)+..., 1/len."""
    v = np.asarray(v, dtype=np.float32)
    acc = np.zeros(v.shape[0], np.float32)
    for k in range(v.shape[1]):
        acc = acc + v[:, k] * v[:, k]
    length = np.sqrt(acc)
    ok = length > np.float32(1.1920929e-7)
    inv = np.where(ok, np.float32(1.0) / np.where(ok, length, np.float32(1.0)), np.float32(0.0)).astype(np.float32)
    out = v * inv[:, None]
    out[~ok] = 0.0
    return out.astype(np.float32)


def unit_dirs(stream: Stream, n: int, dim: int) -> np.ndarray:
    """Uniform unit directions: rejection from [-1,1]^D (keep 1e-6 < |v|^2 <= 1), then Ray::new."""
    out = np.empty((0, dim), np.float32)
    while out.shape[0] < n:
        m = max(1024, int((n - out.shape[0]) * (2.2 if dim <= 3 else 2 ** dim / _ball_vol(dim) * 1.1)))
        c = stream.range(-1.0, 1.0, m * dim).reshape(m, dim)
        r2 = np.zeros(m, np.float32)
        for k in range(dim):
            r2 = r2 + c[:, k] * c[:, k]
        keep = (r2 <= np.float32(1.0)) & (r2 > np.float32(1e-6))
        out = np.concatenate([out, c[keep]])
    return normalize_or_zero(out[:n])


def _ball_vol(d: int) -> float:
    return math.pi ** (d / 2) / math.gamma(d / 2 + 1)


def rays_array(origins: np.ndarray, dirs: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.concatenate([origins, dirs], axis=1), dtype=np.float32)


# ---------------------------------------------------------------------------
# Scenes of the benchmark configurations
# ---------------------------------------------------------------------------

@dataclass
class Scene:
    name: str
    dim: int
    kind: int           # 0 AABB, 1 SPHERE, 2 TRIANGLE
    prims: np.ndarray   # caller layout (include/sq.h)
    max_toi: float
    meta: dict


def boxes_scene(n=10_000, seed=0x5A17_C0DE_0000_0001, lo=-100.0, hi=100.0, hmin=0.1, hmax=2.0, dim=3) -> Scene:
    """Axis-aligned boxes: centres U[lo,hi]^D, half-extents U[hmin,hmax] per axis.
    Per box the stream yields D centre draws then D half-extent draws."""
    s = Stream(seed)
    d = s.u64(n * 2 * dim).reshape(n, 2 * dim)
    c = frange(d[:, :dim], lo, hi)
    h = frange(d[:, dim:], hmin, hmax)
    prims = np.concatenate([c - h, c + h], axis=1).astype(np.float32)
    return Scene("boxes", dim, 0, prims, 400.0, dict(seed=seed, n=n))


def random_rays(n, seed, lo, hi, dim=3) -> np.ndarray:
    """Origins U[lo,hi]^D (one stream), unit directions (second stream)."""
    so = Stream(seed)
    o = so.range(lo, hi, n * dim).reshape(n, dim)
    d = unit_dirs(Stream(seed ^ 0xD1B54A32D192ED03), n, dim)
    return rays_array(o, d)


def c0(n_rays=65_536):
    """10,000 boxes in [-100,100]^3, 65,536 rays from U[-150,150]^3, t_max 400."""
    sc = boxes_scene()
    rays = random_rays(n_rays, 0xC0C0_0000_0000_0002, -150.0, 150.0)
    return sc, rays


# --- heightfield -----------------------------------------------------------

def heightfield_params(seed=0x4E16_47F1_E1D0_0003):
    """8 sines: amplitude U[1,40], wavelength log-U[8,512], direction and phase U[0,2pi)."""
    s = Stream(seed)
    a = s.range(1.0, 40.0, 8).astype(np.float64)
    lam = 8.0 * 64.0 ** s.range(0.0, 1.0, 8).astype(np.float64)
    th = s.range(0.0, 2 * math.pi, 8).astype(np.float64)
    ph = s.range(0.0, 2 * math.pi, 8).astype(np.float64)
    return a, lam, th, ph


def height(x, y, params):
    a, lam, th, ph = params
    x = np.asarray(x, np.float64)[..., None]
    y = np.asarray(y, np.float64)[..., None]
    return np.sum(a * np.sin(2 * math.pi * (x * np.cos(th) + y * np.sin(th)) / lam + ph), axis=-1)


def heightfield_scene(grid=1025, extent=512.0, seed=0x4E16_47F1_E1D0_0003) -> Scene:
    """grid x grid vertices over [-extent, extent]^2 (unit spacing for 1025),
    two triangles per cell: (v00, v10, v11) and (v00, v11, v01); cell (i, j)
    (i along x) gives triangles 2*(j*(grid-1)+i) and +1."""
    params = heightfield_params(seed)
    g = grid
    xs = np.linspace(-extent, extent, g).astype(np.float32)
    X, Y = np.meshgrid(xs, xs, indexing="xy")            # [j, i]
    Z = height(X.astype(np.float64), Y.astype(np.float64), params).astype(np.float32)
    V = np.stack([X, Y, Z], axis=-1)                      # [j, i, 3]
    v00 = V[:-1, :-1]; v10 = V[:-1, 1:]; v11 = V[1:, 1:]; v01 = V[1:, :-1]
    t0 = np.concatenate([v00, v10, v11], axis=-1)
    t1 = np.concatenate([v00, v11, v01], axis=-1)
    tris = np.stack([t0, t1], axis=2).reshape(-1, 9)
    return Scene("heightfield", 3, 2, np.ascontiguousarray(tris, np.float32), 2000.0,
                 dict(seed=seed, grid=grid, zmin=float(Z.min()), zmax=float(Z.max()), extent=extent, params=params))


def camera_rays(sc: Scene, res=2048, seed=0xCA3E_7A00_0000_0004, fov_deg=60.0, jitter_seed=None) -> np.ndarray:
    """Pinhole camera at a seeded pose 200 units above the surface (scaled by
    extent / 512 for smaller heightfields), looking
    down at a seeded pitch in [-50, -20] degrees; one ray per pixel centre, or,
    with jitter_seed, per uniformly drawn point of each pixel."""
    s = Stream(seed)
    e = 0.5 * sc.meta["extent"]
    cx, cy = (float(v) for v in s.range(-e, e, 2))
    yaw = float(s.range(0.0, 2 * math.pi, 1)[0])
    pitch = math.radians(float(s.range(-50.0, -20.0, 1)[0]))
    cz = float(height(cx, cy, sc.meta["params"])) + 200.0 * sc.meta["extent"] / 512.0
    fwd = np.array([math.cos(pitch) * math.cos(yaw), math.cos(pitch) * math.sin(yaw), math.sin(pitch)])
    right = np.array([math.sin(yaw), -math.cos(yaw), 0.0])
    up = np.cross(right, fwd)
    half = math.tan(math.radians(fov_deg) / 2)
    px = (np.arange(res, dtype=np.float64) + 0.5) / res * 2 - 1
    U, Vv = np.meshgrid(px * half, -px * half, indexing="xy")
    if jitter_seed is not None:
        j = Stream(jitter_seed).range(-0.5, 0.5, 2 * res * res).reshape(2, res, res).astype(np.float64)
        U = U + j[0] * (2.0 / res) * half
        Vv = Vv - j[1] * (2.0 / res) * half
    d = (fwd[None, None, :] + U[..., None] * right[None, None, :] + Vv[..., None] * up[None, None, :]).reshape(-1, 3)
    d = normalize_or_zero(d.astype(np.float32))
    o = np.broadcast_to(np.array([cx, cy, cz], np.float32), d.shape)
    return rays_array(o, d)


def incoherent_rays(sc: Scene, n, seed=0x1AC0_4E2E_0000_0005) -> np.ndarray:
    """Origins uniform in the scene's bounding box, directions uniform on the sphere."""
    e = sc.meta["extent"]
    s = Stream(seed)
    d = s.u64(n * 3).reshape(n, 3)
    o = np.stack([frange(d[:, 0], -e, e), frange(d[:, 1], -e, e), frange(d[:, 2], sc.meta["zmin"], sc.meta["zmax"])], axis=1)
    dirs = unit_dirs(Stream(seed ^ 0xD1B54A32D192ED03), n, 3)
    return rays_array(o.astype(np.float32), dirs)


def c1(res=2048, n_incoherent=None, grid=1025, extent=512.0):
    """2,097,152-triangle heightfield; 4M coherent camera rays + 4M incoherent rays; t_max 2000."""
    sc = heightfield_scene(grid=grid, extent=extent)
    coh = camera_rays(sc, res=res)
    inc = incoherent_rays(sc, res * res if n_incoherent is None else n_incoherent)
    return sc, coh, inc


def proposal_training_rays(sc: Scene, n: int, seed=0x7EA1_0000_0000_0008) -> np.ndarray:
    """Training rays for the learned proposal: fresh draws from the evaluation
    distributions.  n/2 camera rays from the same seeded camera as camera_rays
    at points drawn uniformly inside the pixels (the evaluation uses the pixel
    centres), over as many jittered frames as needed, and n/2 incoherent rays
    from another seed."""
    s = Stream(seed)
    half = n // 2
    res = max(1, int(math.sqrt(half)))
    cams, got = [], 0
    while got < half:
        r = camera_rays(sc, res=res, jitter_seed=int(s.u64(1)[0]))
        cams.append(r)
        got += r.shape[0]
    coh = np.concatenate(cams)[:half]
    inc = incoherent_rays(sc, n - half, seed=int(s.u64(1)[0]))
    return np.ascontiguousarray(np.concatenate([coh, inc]), np.float32)


# --- D-dimensional boxes from a Gaussian mixture ----------------------------

def gmm_boxes_scene(n=1_000_000, dim=6, k=32, seed=0x6A55_1A40_0000_0006) -> Scene:
    """Centres from a k-component mixture (means U[-500,500]^D, sigma 5),
    half-extents U[0.05, 1].  Normals by Box-Muller in float64."""
    s = Stream(seed)
    means = s.range(-500.0, 500.0, k * dim).reshape(k, dim).astype(np.float64)
    comp = (s.u64(n) % np.uint64(k)).astype(np.int64)
    u1 = s.range(0.0, 1.0, n * dim).astype(np.float64).reshape(n, dim)
    u2 = s.range(0.0, 1.0, n * dim).astype(np.float64).reshape(n, dim)
    z = np.sqrt(-2.0 * np.log(np.maximum(u1, 1e-12))) * np.cos(2 * math.pi * u2)
    c = (means[comp] + 5.0 * z).astype(np.float32)
    h = s.range(0.05, 1.0, n * dim).reshape(n, dim)
    prims = np.concatenate([c - h, c + h], axis=1).astype(np.float32)
    return Scene("gmm_boxes", dim, 0, prims, 1000.0, dict(seed=seed, means=means))


def gmm_points(sc: Scene, n: int, seed: int, sigma=10.0) -> np.ndarray:
    """Points near the mixture's components (a random mean + N(0, sigma^2)
    per axis): query origins and centres that land where the boxes are."""
    means = sc.meta["means"]
    dim = means.shape[1]
    s = Stream(seed)
    comp = (s.u64(n) % np.uint64(means.shape[0])).astype(np.int64)
    u1 = s.range(0.0, 1.0, n * dim).astype(np.float64).reshape(n, dim)
    u2 = s.range(0.0, 1.0, n * dim).astype(np.float64).reshape(n, dim)
    z = np.sqrt(-2.0 * np.log(np.maximum(u1, 1e-12))) * np.cos(2 * math.pi * u2)
    return (means[comp] + sigma * z).astype(np.float32)


def c4_queries(sc: Scene, n_rays=1_000_000, n_overlap=262_144, n_cast=262_144, seed=0xC4C4_0000_0009):
    """C4 queries on the D-dimensional mixture scene: rays from points near the
    components with uniform directions; overlap boxes centred near the
    components with half-extents U[1, 8]; box casts (half-extent 0.5) from
    points near the components, uniform directions, sweep length U(0, 50]."""
    dim = sc.dim
    o = gmm_points(sc, n_rays, seed)
    rays = rays_array(o, unit_dirs(Stream(seed + 1), n_rays, dim))
    c = gmm_points(sc, n_overlap, seed + 2)
    h = Stream(seed + 3).range(1.0, 8.0, n_overlap * dim).reshape(n_overlap, dim)
    regions = np.concatenate([c - h, c + h], axis=1).astype(np.float32)
    co = gmm_points(sc, n_cast, seed + 4)
    cd = unit_dirs(Stream(seed + 5), n_cast, dim)
    half = np.full((n_cast, dim), 0.5, np.float32)
    casts = np.concatenate([co, cd, -half, half], axis=1).astype(np.float32)
    lim = (50.0 - Stream(seed + 6).range(0.0, 50.0, n_cast)).astype(np.float32)   # (0, 50]
    return rays, regions, casts, lim


# --- instanced meshes --------------------------------------------------------

_ICO_V = None


def _icosahedron():
    global _ICO_V
    if _ICO_V is None:
        p = (1 + 5 ** 0.5) / 2
        v = np.array([[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0], [0, -1, p], [0, 1, p], [0, -1, -p], [0, 1, -p],
                      [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]], np.float64)
        f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                      [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                      [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
        _ICO_V = (v / np.linalg.norm(v, axis=1, keepdims=True), f)
    return _ICO_V


def noisy_sphere(freq: int, seed: int, amp=0.15) -> np.ndarray:
    """Geodesic sphere: each icosahedron face split into freq^2 triangles,
    projected to the unit sphere and displaced radially by a seeded sum of
    three sines of the direction.  Returns [20 freq^2, 9] f32 triangles."""
    v, f = _icosahedron()
    s = Stream(seed)
    k = s.range(1.0, 6.0, 9).astype(np.float64).reshape(3, 3)
    ph = s.range(0.0, 2 * math.pi, 3).astype(np.float64)
    tris = []
    n = freq
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    up = (ii + jj) < n                      # upward triangles
    dn = (ii + jj) < n - 1                  # downward triangles
    for a, b, c in f:
        A, B, Cc = v[a], v[b], v[c]

        def P(i, j):
            w = np.stack([i / n, j / n], -1)
            return A + w[..., :1] * (B - A) + w[..., 1:] * (Cc - A)
        i0, j0 = ii[up], jj[up]
        tris.append(np.stack([P(i0, j0), P(i0 + 1, j0), P(i0, j0 + 1)], 1))
        i1, j1 = ii[dn], jj[dn]
        tris.append(np.stack([P(i1 + 1, j1), P(i1 + 1, j1 + 1), P(i1, j1 + 1)], 1))
    t = np.concatenate(tris)                # [T, 3, 3]
    d = t / np.linalg.norm(t, axis=-1, keepdims=True)
    r = 1.0 + amp * np.sum(np.sin(d @ k.T + ph), axis=-1, keepdims=True) / 3.0
    return np.ascontiguousarray((d * r).reshape(-1, 9), np.float32)


def random_rotations(stream: Stream, n: int) -> np.ndarray:
    """Uniform random rotations (Shoemake quaternions) as row-major f32 3x3."""
    u = stream.range(0.0, 1.0, n * 3).astype(np.float64).reshape(n, 3)
    a, b = np.sqrt(1 - u[:, 0]), np.sqrt(u[:, 0])
    x, y = a * np.sin(2 * math.pi * u[:, 1]), a * np.cos(2 * math.pi * u[:, 1])
    z, w = b * np.sin(2 * math.pi * u[:, 2]), b * np.cos(2 * math.pi * u[:, 2])
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                  2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                  2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1)
    return np.ascontiguousarray(R.reshape(n, 3, 3), np.float32)


def instanced_scene(n_meshes=16, n_inst=4096, tri_lo=2000, tri_hi=50000, extent=1000.0, seed=0x2C2C_0000_0007):
    """Meshes with triangle counts log-uniform in [tri_lo, tri_hi] (rounded to
    20 f^2), instances with uniform rotations, translations U[-extent, extent]^3
    and uniform scales U[0.5, 4]."""
    s = Stream(seed)
    counts = np.exp(s.range(math.log(tri_lo), math.log(tri_hi), n_meshes).astype(np.float64))
    freqs = np.maximum(1, np.round(np.sqrt(counts / 20.0))).astype(int)
    meshes = [noisy_sphere(int(fq), seed + 101 * (m + 1)) for m, fq in enumerate(freqs)]
    mesh_of = (s.u64(n_inst) % np.uint64(n_meshes)).astype(np.uint32)
    rot = random_rotations(s, n_inst)
    trans = s.range(-extent, extent, n_inst * 3).reshape(n_inst, 3)
    scale = s.range(0.5, 4.0, n_inst)
    return dict(meshes=meshes, mesh_of=mesh_of, rot=rot, trans=trans, scale=scale)
