"""Pins of the f64 oracle (oracle/sq_oracle64.cpp: the reference's `f64`
feature, sq/src/maths/point.rs:14-18), on the CPU:

  * closed forms: slab entries and face normals of axis-aligned rays, the ball's
    entry root from outside and exit root from inside, a tangent ray, a miss;
  * dyadic scenes, where every slab and centroid operation is exact in both
    precisions: the f64 naive scans equal the f32 oracle's BVH walk bit for bit
    (nearest ids and toi, all hits, overlap, filters) and the f64 tree equals
    the f32 tree (node count, leaves, depth, surface sum);
  * random balls against an independent float128-free root solver (numpy
    float64 with the cancellation-free quadratic form): same nearest ids where
    the two nearest candidates are separated, toi within 1e-12 relative, and
    every nearest toi is a root (|at(t) - c| = r);
  * all hits: sorted by (toi, leaf), and the first hit is the nearest;
  * the consumer kind (a test_ray in double through a C function): the
    built-in ball's arithmetic as a consumer equals the SPHERE kind bit for bit
    in every walk and mode; with shrunk boxes the reference's walk and the
    naive scan differ, every answer is a true hit, and the guarded forms agree.
"""
import numpy as np
import pytest

import oracle


def test_closed_forms():
    box = np.array([[0.0, 0.0, 0.0, 1.0, 1.0, 1.0]])
    s = oracle.Scene64(3, oracle.BOX, box)
    rays = np.array([[-5.0, 0.5, 0.5, 1.0, 0.0, 0.0],
                     [0.5, 0.5, 7.0, 0.0, 0.0, -1.0],
                     [0.5, 0.5, 0.5, 0.0, 1.0, 0.0],    # inside: entry 0, no entry axis
                     [-5.0, 3.0, 0.5, 1.0, 0.0, 0.0]])  # parallel slab outside: miss
    w = s.raycast_nearest(rays, 100.0)
    assert list(w["leaf"]) == [0, 0, 0, -1]
    assert list(w["t"][:3]) == [5.0, 6.0, 0.0] and np.isinf(w["t"][3])
    assert np.array_equal(w["normal"][0], [-1.0, 0.0, 0.0])
    assert np.array_equal(w["normal"][1], [0.0, 0.0, 1.0])
    assert np.array_equal(w["normal"][2], [0.0, 0.0, 0.0])
    assert np.array_equal(w["point"][1], [0.5, 0.5, 1.0])
    # max_toi cuts a hit at 5 but not at exactly 5
    assert s.raycast_nearest(rays[:1], 4.999)["leaf"][0] == -1
    assert s.raycast_nearest(rays[:1], 5.0)["leaf"][0] == 0

    ball = np.array([[10.0, 0.0, 0.0, 2.0]])
    b = oracle.Scene64(3, oracle.SPHERE, ball)
    rays = np.array([[0.0, 0.0, 0.0, 1.0, 0.0, 0.0],     # entry root 8
                     [10.0, 0.0, 0.0, 0.0, 1.0, 0.0],    # from the centre: exit root 2
                     [0.0, 2.0, 0.0, 1.0, 0.0, 0.0],     # tangent at (10, 2, 0): t = 10
                     [0.0, 2.5, 0.0, 1.0, 0.0, 0.0],     # miss
                     [20.0, 0.0, 0.0, 1.0, 0.0, 0.0]])   # behind the ray: miss
    w = b.raycast_nearest(rays, 100.0)
    assert list(w["leaf"]) == [0, 0, 0, -1, -1]
    assert list(w["t"][:3]) == [8.0, 2.0, 10.0]
    assert np.array_equal(w["normal"][0], [-1.0, 0.0, 0.0])
    assert np.array_equal(w["normal"][1], [0.0, 1.0, 0.0])
    # overlap: the ball's nearest-point test, not its box
    regions = np.array([[11.5, 1.5, -0.1, 12.0, 2.0, 0.1],    # box corner region outside the ball
                        [11.0, 1.0, -0.1, 12.0, 2.0, 0.1],    # distance sqrt(2) < 2
                        [12.0, -1.0, -1.0, 13.0, 1.0, 1.0]])  # touching at x = 12
    off, leaf = b.overlap(regions)
    assert list(off) == [0, 0, 1, 2] and list(leaf) == [0, 0]


def dyadic_boxes(n, dim, seed, span):
    rng = np.random.default_rng(seed)
    lo = rng.integers(-span, span, (n, dim)) / 8.0
    ext = rng.integers(1, 24, (n, dim)) / 8.0
    return np.concatenate([lo, lo + ext], 1)


def dyadic_rays(n, dim, seed, span, pad, zero_dirs=True):
    rng = np.random.default_rng(seed)
    o = rng.integers(-span - pad, span + pad, (n, dim)) / 8.0
    d = rng.choice([-1.0, -0.5, -0.25, 0.0, 0.25, 0.5, 1.0] if zero_dirs else [-1.0, -0.5, -0.25, 0.25, 0.5, 1.0], (n, dim))
    d[np.all(d == 0, axis=1), 0] = 1.0
    return np.concatenate([o, d], 1)


@pytest.mark.parametrize("dim", [2, 3, 5])
def test_dyadic_box_scenes_equal_the_f32_oracle(dim):
    span, pad = {2: (256, 64), 3: (256, 64), 5: (24, 16)}[dim]   # denser in 5-D, so rays still hit
    boxes = dyadic_boxes(3000, dim, 11 + dim, span)
    rays = dyadic_rays(3000, dim, 21 + dim, span, pad, zero_dirs=dim < 5)
    layers = np.random.default_rng(5).integers(0, 4, 3000).astype(np.uint64)
    s64 = oracle.Scene64(dim, oracle.BOX, boxes, layers=layers)
    s32 = oracle.Scene(dim, oracle.BOX, boxes.astype(np.float32), layers=layers)
    for mask, exclude in [(oracle.ALL_LAYERS, -1), (0b10, 17)]:
        w64 = s64.raycast_nearest(rays, 50.0, mask=mask, exclude=exclude)
        w32 = s32.raycast_nearest(rays.astype(np.float32), 50.0, mask=mask, exclude=exclude)
        assert (w32["leaf"] >= 0).sum() > 500
        assert np.array_equal(w64["leaf"], w32["leaf"])
        assert np.array_equal(w64["t"], w32["t"].astype(np.float64))
        hit = w32["leaf"] >= 0
        assert np.array_equal(w64["normal"][hit], w32["normal"][hit].astype(np.float64))
    a64 = s64.raycast_all(rays[:800], 50.0)
    a32 = s32.raycast_all(rays[:800].astype(np.float32), 50.0)
    assert np.array_equal(a64[0], a32[0].astype(np.uint64)) and np.array_equal(a64[1], a32[1])
    assert np.array_equal(a64[2], a32[2].astype(np.float64))
    regions = np.concatenate([boxes[:500, :dim] - 1.0, boxes[:500, :dim] + 0.5], 1)
    o64 = s64.overlap(regions, mask=0b1)
    o32 = s32.overlap(regions.astype(np.float32), mask=0b1)
    assert np.array_equal(o64[0], o32[0].astype(np.uint64)) and np.array_equal(o64[1], o32[1])
    # the f64 median split builds the f32 tree when every centroid is exact
    t64, t32 = s64.tree_stats(), s32.tree_stats()
    for k in ("node_count", "leaf_count", "prim_count", "max_depth"):
        assert t64[k] == t32[k], k
    # same boxes, so the same surface sum, up to the f32 sum's rounding
    assert t64["quality"] == pytest.approx(float(t32["quality"]), rel=1e-5)
    assert t64["bytes"] == t64["node_count"] * (16 * dim + 16) + 3000 * 4


def independent_ball_roots(o, d, c, r):
    """Both roots of |o + t d - c| = r by the cancellation-free form (not the
    oracle's -b -+ sqrt(disc)), nan where the ray misses."""
    oc = o - c
    a = np.sum(d * d, -1)
    b = np.sum(oc * d, -1)
    cc = np.sum(oc * oc, -1) - r * r
    disc = b * b - a * cc
    sq = np.sqrt(np.where(disc >= 0, disc, np.nan))
    q = -(b + np.copysign(sq, b))
    t0, t1 = q / a, cc / q
    return np.minimum(t0, t1), np.maximum(t0, t1)


def test_random_balls_against_an_independent_solver():
    rng = np.random.default_rng(7)
    n, nr = 600, 1500
    balls = np.concatenate([rng.uniform(-50, 50, (n, 3)), rng.uniform(0.5, 3.0, (n, 1))], 1)
    o = rng.uniform(-60, 60, (nr, 3))
    d = rng.normal(size=(nr, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], 1)
    w = oracle.Scene64(3, oracle.SPHERE, balls).raycast_nearest(rays, 500.0)
    lo, hi = independent_ball_roots(o[:, None, :], d[:, None, :], balls[None, :, :3], balls[None, :, 3])
    t = np.where(lo >= 0, lo, hi)
    t = np.where((t >= 0) & (t <= 500.0), t, np.inf)
    best = np.argmin(t, 1)
    tb = t[np.arange(nr), best]
    srt = np.sort(t, 1)
    with np.errstate(invalid="ignore"):
        separated = (srt[:, 1] - srt[:, 0]) > 1e-9 * np.maximum(1.0, srt[:, 0])
    hit = np.isfinite(tb)
    assert hit.sum() > 300
    ok = separated & hit
    assert np.array_equal(w["leaf"][ok], best[ok])
    assert np.allclose(w["t"][ok], tb[ok], rtol=1e-12, atol=0)
    assert (w["leaf"][~hit] == -1).all()
    h = w["leaf"] >= 0
    c = balls[w["leaf"][h]]
    p = o[h] + d[h] * w["t"][h][:, None]
    assert np.allclose(np.linalg.norm(p - c[:, :3], axis=1), c[:, 3], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("kind", [oracle.BOX, oracle.SPHERE])
def test_all_hits_sorted_and_headed_by_the_nearest(kind):
    rng = np.random.default_rng(9 + kind)
    dim = 4
    if kind == oracle.BOX:
        lo = rng.uniform(-8, 8, (800, dim))
        prims = np.concatenate([lo, lo + rng.uniform(0.5, 3.0, (800, dim))], 1)
    else:
        prims = np.concatenate([rng.uniform(-8, 8, (800, dim)), rng.uniform(0.5, 2.0, (800, 1))], 1)
    rays = np.concatenate([rng.uniform(-8, 8, (700, dim)), rng.normal(size=(700, dim))], 1)
    # guarded leaf tests: the nearest hit is the first of all hits on every ray;
    # the reference's (plain) walk can end on a later one, never an earlier one
    # (a hit outside its ball's box may be pruned once a hit below that box's
    # entry is known), and both find a hit on the same rays
    for guarded in (True, False):
        s = oracle.Scene64(dim, kind, prims, guarded=guarded)
        off, leaf, t = s.raycast_all(rays, 80.0)
        w = s.raycast_nearest(rays, 80.0)
        off = off.astype(np.int64)
        assert off[-1] > 300
        for i in range(700):
            a, b = off[i], off[i + 1]
            if a == b:
                assert w["leaf"][i] == -1
                continue
            keys = list(zip(t[a:b], leaf[a:b]))
            assert keys == sorted(keys)
            if guarded or kind == oracle.BOX:
                assert leaf[a] == w["leaf"][i] and t[a] == w["t"][i]
            else:
                assert w["leaf"][i] >= 0 and (t[a], leaf[a]) <= (w["t"][i], w["leaf"][i])


def test_shape_cast_closed_forms_and_identities():
    # a box probe of half-extent 0.25 swept along +x meets the unit box at 5 - 0.25
    s = oracle.Scene64(3, oracle.BOX, np.array([[0.0, 0.0, 0.0, 1.0, 1.0, 1.0]]))
    cast = np.array([[-5.0, 0.5, 0.5, 1.0, 0.0, 0.0, -0.25, -0.25, -0.25, 0.25, 0.25, 0.25]])
    w = s.shapecast_nearest(cast, 100.0)
    assert w["leaf"][0] == 0 and w["t"][0] == 4.75 and np.array_equal(w["point"][0], [-0.25, 0.5, 0.5])
    # a ball probe (radius probe_max[0]) against a ball: contact at the summed radius
    b = oracle.Scene64(3, oracle.SPHERE, np.array([[10.0, 0.0, 0.0, 2.0]]))
    cast = np.array([[0.0, 0.0, 0.0, 1.0, 0.0, 0.0, -0.5, -0.5, -0.5, 0.5, 0.5, 0.5]])
    assert b.shapecast_nearest(cast, 100.0)["t"][0] == 7.5


@pytest.mark.parametrize("kind", [oracle.BOX, oracle.SPHERE])
def test_dyadic_shape_casts(kind):
    """Dyadic scenes: BOX casts equal the f32 oracle's BVH casts bit for bit; a
    SPHERE cast with probe half-extent h equals the ray cast against balls of
    radius r + h (every sum, grown box and product is exact)."""
    dim = 3
    rng = np.random.default_rng(31 + kind)
    if kind == oracle.BOX:
        prims = dyadic_boxes(3000, dim, 41, 256)
    else:
        prims = np.concatenate([rng.integers(-256, 256, (3000, dim)) / 8.0, rng.integers(2, 24, (3000, 1)) / 8.0], 1)
    rays = dyadic_rays(2000, dim, 43 + kind, 256, 64)
    h = rng.integers(1, 8, (2000, 1)) / 8.0
    casts = np.concatenate([rays, np.repeat(-h, dim, 1), np.repeat(h, dim, 1)], 1)
    s64 = oracle.Scene64(dim, kind, prims)
    w = s64.shapecast_nearest(casts, 60.0)
    assert (w["leaf"] >= 0).sum() > 300
    a64 = s64.shapecast_all(casts[:600], 60.0)
    if kind == oracle.BOX:
        s32 = oracle.Scene(dim, oracle.BOX, prims.astype(np.float32))
        w32 = s32.shapecast_nearest(casts.astype(np.float32), 60.0)
        assert np.array_equal(w["leaf"], w32["leaf"]) and np.array_equal(w["t"], w32["t"].astype(np.float64))
        a32 = s32.shapecast_all(casts[:600].astype(np.float32), 60.0)
        assert np.array_equal(a64[0], a32[0].astype(np.uint64)) and np.array_equal(a64[1], a32[1])
        assert np.array_equal(a64[2], a32[2].astype(np.float64))
    else:
        for hv in np.unique(h):
            sel = h[:, 0] == hv
            grown = prims.copy()
            grown[:, dim] += hv
            wr = oracle.Scene64(dim, kind, grown).raycast_nearest(rays[sel], 60.0)
            assert np.array_equal(w["leaf"][sel], wr["leaf"]) and np.array_equal(w["t"][sel], wr["t"])


def test_per_query_limits_equal_scalar_calls():
    rng = np.random.default_rng(55)
    prims = np.concatenate([rng.uniform(-20, 20, (1500, 3)), rng.uniform(0.3, 2.0, (1500, 1))], 1)
    rays = np.concatenate([rng.uniform(-25, 25, (2000, 3)), rng.normal(size=(2000, 3))], 1)
    casts = np.concatenate([rays, np.full((2000, 3), -0.3), np.full((2000, 3), 0.3)], 1)
    lim = rng.choice([3.0, 12.0, 40.0], 2000)
    s = oracle.Scene64(3, oracle.SPHERE, prims)
    w = s.raycast_nearest(rays, lim)
    c = s.shapecast_nearest(casts, lim)
    a = s.shapecast_all(casts, lim)
    off = a[0].astype(np.int64)
    for v in (3.0, 12.0, 40.0):
        sel = np.flatnonzero(lim == v)
        ws = s.raycast_nearest(rays[sel], v)
        assert np.array_equal(w["leaf"][sel], ws["leaf"]) and np.array_equal(w["t"][sel], ws["t"])
        cs = s.shapecast_nearest(casts[sel], v)
        assert np.array_equal(c["leaf"][sel], cs["leaf"]) and np.array_equal(c["t"][sel], cs["t"])
        as_ = s.shapecast_all(casts[sel], v)
        got = np.concatenate([a[1][off[i]:off[i + 1]] for i in sel])
        assert np.array_equal(got, as_[1])
    assert (w["leaf"] >= 0).sum() > 300 and (w["t"][w["leaf"] >= 0] <= lim[w["leaf"] >= 0]).all()


@pytest.mark.parametrize("dim,kind", [(2, oracle.BOX), (3, oracle.SPHERE), (5, oracle.BOX), (6, oracle.SPHERE)])
def test_bvh_walk_equals_naive(dim, kind):
    """The reference's contract (sq/src/lib.rs:15-17): the BVH answers what the
    naive scan does, here in double over the f64 tree; and the parallel batch
    equals the sequential one."""
    rng = np.random.default_rng(70 + dim + kind)
    span = {2: 10.0, 3: 10.0, 5: 3.0, 6: 2.5}[dim]   # denser in more dimensions, so rays still hit
    c = rng.uniform(-span, span, (2500, dim))
    if kind == oracle.BOX:
        prims = np.concatenate([c, c + rng.uniform(0.2, 2.0, (2500, dim))], 1)
    else:
        prims = np.concatenate([c, rng.uniform(0.2, 1.5, (2500, 1))], 1)
    rays = np.concatenate([rng.uniform(-span, span, (3000, dim)), rng.normal(size=(3000, dim))], 1)
    s = oracle.Scene64(dim, kind, prims)
    want = s.raycast_nearest(rays, 50.0)
    got = s.bvh_raycast_nearest(rays, 50.0, threads=1)
    par = s.bvh_raycast_nearest(rays, 50.0, threads=7)
    assert (want["leaf"] >= 0).sum() > 500
    assert np.array_equal(got["leaf"], want["leaf"]) and np.array_equal(got["t"], want["t"])
    assert np.array_equal(par["leaf"], got["leaf"]) and np.array_equal(par["t"], got["t"])
    assert par["nodes_tested"] == got["nodes_tested"] and par["prims_tested"] == got["prims_tested"]
    assert got["nodes_tested"] < 3000 * s.tree_stats()["node_count"] // 4   # the tree prunes


@pytest.mark.parametrize("dim", [2, 3, 5])
def test_box_instances_dyadic_equal_the_f32_oracle(dim):
    """Box instances (two_level.rs:70-167) in double: with signed-permutation
    rotations and dyadic translations, half extents and rays every operation is
    exact in both precisions, so the f64 scan equals the f32 oracle's guarded
    box instances bit for bit, world boxes included."""
    rng = np.random.default_rng(90 + dim)
    n = 400
    rot = np.zeros((n, dim, dim))
    for i in range(n):
        rot[i, np.arange(dim), rng.permutation(dim)] = rng.choice([-1.0, 1.0], dim)
    span = {2: 160, 3: 160, 5: 24}[dim]   # denser in 5-D, so rays still hit
    trans = rng.integers(-span, span, (n, dim)) / 8.0
    halves = rng.integers(1, 24, (n, dim)) / 8.0
    rays = dyadic_rays(3000, dim, 95 + dim, span, 16, zero_dirs=False)
    s64 = oracle.BoxInstances64(dim, rot, trans, halves)
    w = s64.raycast_nearest(rays, 80.0, world=True)
    s32 = oracle.InstanceScene(dim, oracle.INST_BOX_GUARDED, rot.astype(np.float32), trans.astype(np.float32),
                               halves=halves.astype(np.float32))
    w32 = s32.raycast_nearest(rays.astype(np.float32), 80.0, naive=True)
    assert (w["leaf"] >= 0).sum() > 300
    assert np.array_equal(w["leaf"], w32["leaf"]) and np.array_equal(w["t"], w32["t"].astype(np.float64))
    for i in range(0, n, 37):
        box = np.concatenate([-halves[i], halves[i]]).astype(np.float32)
        wb = oracle.iso_transform_aabb(rot[i].astype(np.float32), trans[i].astype(np.float32), box)
        assert np.array_equal(w["world"][i], np.asarray(wb, np.float64).reshape(-1))


# ---------------------------------------------------------------------------
# The f64 Tlas queries over box and ball instances (oracle.Instances64)
# ---------------------------------------------------------------------------

def _rays64(rng, n, span, dim=3):
    o = rng.uniform(-span, span, (n, dim))
    d = rng.normal(size=(n, dim))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1)


def _casts64(rays, half):
    n, dim = rays.shape[0], rays.shape[1] // 2
    return np.concatenate([rays, np.full((n, dim), -half), np.full((n, dim), half)], 1)


def test_ball_instance_closed_forms():
    """A ball instance of radius 1 at (5, 0, 0): a ray along +x from the origin
    enters at 4, one from the centre leaves at 1, one passing at height 2
    misses; a cube probe of half 0.5 swept along +x touches at 5 - 1.5 = 3.5
    (shape_overlap.rs:326-337); a region whose nearest point lies 0.8 from the
    centre overlaps, one that touches the ball's world box but lies
    sqrt(3 * 0.8^2) = 1.39 away does not (point_to_aabb, :59-72); a box
    instance answers no cast (the trait default) and overlaps at AABB level:
    every instance of a leaf node touching the region, even one whose own
    world box misses it (a 3-instance Tlas is one leaf)."""
    ball = oracle.Instances64(3, 1, np.eye(3)[None], np.array([[5.0, 0.0, 0.0]]), np.array([1.0]))
    rays = np.array([[0, 0, 0, 1, 0, 0], [5, 0, 0, 1, 0, 0], [0, 2, 0, 1, 0, 0]], np.float64)
    r = ball.raycast_nearest(rays, 100.0)
    assert r["leaf"].tolist() == [0, 0, -1] and r["t"][:2].tolist() == [4.0, 1.0]
    off, leaf, t = ball.raycast_all(rays, 100.0)
    assert off.tolist() == [0, 1, 2, 2] and t.tolist() == [4.0, 1.0]
    c = ball.shapecast_nearest(_casts64(rays[:1], 0.5), 100.0)
    assert c["leaf"][0] == 0 and c["t"][0] == 3.5
    regions = np.array([[3.5, -1, -1, 4.2, 1, 1], [5.8, 0.8, 0.8, 6.5, 1.5, 1.5]], np.float64)
    off, leaf = ball.overlap(regions)
    assert off.tolist() == [0, 1, 1] and leaf.tolist() == [0]
    boxes = oracle.Instances64(3, 0, np.tile(np.eye(3), (3, 1, 1)), np.array([[5.0, 0, 0], [9.0, 0, 0], [30.0, 0, 0]]),
                               np.ones((3, 3)))
    assert boxes.raycast_nearest(rays[:1], 100.0)["t"][0] == 4.0
    off, leaf, t = boxes.raycast_all(rays[:1], 100.0)
    assert leaf.tolist() == [0, 1, 2] and t.tolist() == [4.0, 8.0, 29.0]
    assert (boxes.shapecast_nearest(_casts64(rays, 0.5), 100.0)["leaf"] == -1).all()
    assert boxes.shapecast_all(_casts64(rays, 0.5), 100.0)[0][-1] == 0
    # the region touches the leaf's box (x 4 .. 31) near x 10 but no instance's own box
    assert boxes.overlap(np.array([[10.5, -0.5, -0.5, 11.0, 0.5, 0.5]]))[1].tolist() == [0, 1, 2]
    assert boxes.overlap(np.array([[10.5, 2.5, -0.5, 11.0, 3.0, 0.5]]))[1].tolist() == []
    assert boxes.overlap(np.array([[9.5, -0.5, -0.5, 11.0, 0.5, 0.5]]))[1].tolist() == [0, 1, 2]
    naive = oracle.Instances64(3, 0, np.tile(np.eye(3), (3, 1, 1)), np.array([[5.0, 0, 0], [9.0, 0, 0], [30.0, 0, 0]]),
                               np.ones((3, 3)), naive=True)
    assert naive.overlap(np.array([[10.5, -0.5, -0.5, 11.0, 0.5, 0.5]]))[1].tolist() == []
    assert naive.overlap(np.array([[9.5, -0.5, -0.5, 11.0, 0.5, 0.5]]))[1].tolist() == [1]


@pytest.mark.parametrize("seed", [1, 2])
def test_tlas_ball_instances_match_the_naive_balls(seed):
    """shape_overlap.rs:347-411 in double: a Tlas of ball instances
    (translations) answers rays, shape casts (nearest and all) and overlap
    exactly as the naive scan over the same balls as a single-level provider
    (the SPHERE kind: ray_ball, the radius-sum cast, point_to_aabb); with
    rotated instances the Tlas walk equals the naive instance scan."""
    rng = np.random.default_rng(700 + seed)
    hits = casts_hit = overlaps = 0
    for trial in range(40):
        n = 48 if trial % 2 else 64
        trans = rng.uniform(-30.0, 30.0, (n, 3))
        radii = rng.uniform(0.5, 2.0, n)
        inst = oracle.Instances64(3, 1, np.tile(np.eye(3), (n, 1, 1)), trans, radii)
        balls = oracle.Scene64(3, oracle.SPHERE, np.concatenate([trans, radii[:, None]], 1), naive=True)
        rays = _rays64(rng, 64, 50.0)
        # half of them aimed near a ball centre, so that many hit
        aim = trans[rng.integers(0, n, 32)] + rng.uniform(-2.0, 2.0, (32, 3)) - rays[:32, :3]
        rays[:32, 3:] = aim / np.linalg.norm(aim, axis=1, keepdims=True)
        casts = _casts64(rays, 0.75)
        a, b = inst.raycast_nearest(rays, 500.0), balls.raycast_nearest(rays, 500.0)
        assert np.array_equal(a["leaf"], b["leaf"]) and np.array_equal(a["t"], b["t"])
        oa, la, ta = inst.raycast_all(rays, 500.0)
        ob, lb, tb = balls.raycast_all(rays, 500.0)
        assert np.array_equal(oa, ob) and np.array_equal(la, lb) and np.array_equal(ta, tb)
        a, b = inst.shapecast_nearest(casts, 500.0), balls.shapecast_nearest(casts, 500.0)
        assert np.array_equal(a["leaf"], b["leaf"]) and np.array_equal(a["t"], b["t"])
        oa, la, ta = inst.shapecast_all(casts, 500.0)
        ob, lb, tb = balls.shapecast_all(casts, 500.0)
        assert np.array_equal(oa, ob) and np.array_equal(la, lb) and np.array_equal(ta, tb)
        c = rng.uniform(-30.0, 30.0, (32, 3)); h = rng.uniform(1.0, 6.0, (32, 3))
        regions = np.concatenate([c - h, c + h], 1)
        oa, la = inst.overlap(regions)
        ob, lb = balls.overlap(regions)
        assert np.array_equal(oa, ob) and np.array_equal(la, lb)
        hits += int((inst.raycast_nearest(rays, 500.0)["leaf"] >= 0).sum())
        casts_hit += int(ta.size)
        overlaps += int(la.size)
        # rotated instances (a ball is invariant, the local ray is not): Tlas == naive instance scan
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        rot = np.tile(q, (n, 1, 1))
        ri = oracle.Instances64(3, 1, rot, trans, radii)
        rn = oracle.Instances64(3, 1, rot, trans, radii, naive=True)
        a, b = ri.raycast_nearest(rays, 500.0), rn.raycast_nearest(rays, 500.0)
        assert np.array_equal(a["leaf"], b["leaf"]) and np.array_equal(a["t"], b["t"])
        oa, la, ta = ri.raycast_all(rays, 500.0)
        ob, lb, tb = rn.raycast_all(rays, 500.0)
        assert np.array_equal(oa, ob) and np.array_equal(la, lb) and np.array_equal(ta, tb)
    assert hits > 300 and casts_hit > 300 and overlaps > 100


def _plane_rotations(rng, n, dim):
    rot = np.tile(np.eye(dim), (n, 1, 1))
    for i in range(n):
        a, b = rng.choice(dim, 2, replace=False)
        th = rng.uniform(0.0, 2 * np.pi)
        rot[i, a, a] = np.cos(th); rot[i, b, b] = np.cos(th)
        rot[i, a, b] = -np.sin(th); rot[i, b, a] = np.sin(th)
    return rot


@pytest.mark.parametrize("dim", [2, 3, 6])
def test_tlas_box_instance_queries(dim):
    """Box instances (two_level.rs:70-174) in double, plane rotations: the Tlas
    walks equal the naive instance scans for all hits (= crossings, one per
    instance) in both modes and for nearest hits in the guarded mode; no
    shape cast ever touches a box instance; overlap holds every instance whose
    world box touches the region and, beyond those, only instances of leaf
    nodes touching it (tlas.rs:393-432 with test_overlap's default)."""
    rng = np.random.default_rng(800 + dim)
    n = 300
    span = {2: 40.0, 3: 12.0, 6: 4.0}[dim]   # denser in higher dimension, so rays still hit
    trans = rng.uniform(-span, span, (n, dim))
    halves = rng.uniform(0.5, 2.5, (n, dim))
    rot = _plane_rotations(rng, n, dim)
    rays = _rays64(rng, 2000, 1.5 * span, dim)
    total = 0
    for guarded in (False, True):
        t = oracle.Instances64(dim, 0, rot, trans, halves, guarded=guarded)
        nv = oracle.Instances64(dim, 0, rot, trans, halves, guarded=guarded, naive=True)
        oa, la, ta = t.raycast_all(rays, 500.0)
        ob, lb, tb = nv.raycast_all(rays, 500.0)
        assert np.array_equal(oa, ob) and np.array_equal(la, lb) and np.array_equal(ta, tb)
        total += la.size
        if guarded:
            a, b = t.raycast_nearest(rays, 500.0), nv.raycast_nearest(rays, 500.0)
            assert np.array_equal(a["leaf"], b["leaf"]) and np.array_equal(a["t"], b["t"])
        casts = _casts64(rays, 0.5)
        assert (t.shapecast_nearest(casts, 500.0)["leaf"] == -1).all() and t.shapecast_all(casts, 500.0)[0][-1] == 0
    assert total > 1000
    c = rng.uniform(-span, span, (400, dim)); h = rng.uniform(0.1, 0.1 * span, (400, dim))
    regions = np.concatenate([c - h, c + h], 1)
    t = oracle.Instances64(dim, 0, rot, trans, halves)
    nv = oracle.Instances64(dim, 0, rot, trans, halves, naive=True)
    oa, la = t.overlap(regions)
    ob, lb = nv.overlap(regions)
    extra = 0
    for i in range(len(regions)):
        got, want = set(la[oa[i]:oa[i + 1]].tolist()), set(lb[ob[i]:ob[i + 1]].tolist())
        assert want <= got
        extra += len(got - want)
    assert lb.size > 200 and extra > 0   # leaf-level reports do occur


@pytest.mark.parametrize("dim", [2, 3])
def test_box_instance_queries_dyadic_equal_the_f32_oracle(dim):
    """With signed-permutation rotations and dyadic translations, half extents
    and rays every operation is exact in both precisions and the two trees
    are the same, so the f64 Tlas walks equal the f32 oracle's Tlas (plain box
    instances) bit for bit: all hits and overlap (leaf-level sets included)."""
    rng = np.random.default_rng(900 + dim)
    n = 300
    rot = np.zeros((n, dim, dim))
    for i in range(n):
        rot[i, np.arange(dim), rng.permutation(dim)] = rng.choice([-1.0, 1.0], dim)
    trans = rng.integers(-160, 160, (n, dim)) / 8.0
    halves = rng.integers(1, 24, (n, dim)) / 8.0
    rays = dyadic_rays(2000, dim, 905 + dim, 160, 16, zero_dirs=False)
    s64 = oracle.Instances64(dim, 0, rot, trans, halves)
    s32 = oracle.InstanceScene(dim, oracle.INST_BOX, rot.astype(np.float32), trans.astype(np.float32),
                               halves=halves.astype(np.float32))
    o64, l64, t64 = s64.raycast_all(rays, 80.0)
    o32, l32, t32, _ = s32.raycast_all(rays.astype(np.float32), 80.0)
    assert l64.size > 500
    assert np.array_equal(o64, o32) and np.array_equal(l64, l32) and np.array_equal(t64, t32.astype(np.float64))
    c = rng.integers(-160, 160, (300, dim)) / 8.0; h = rng.integers(1, 40, (300, dim)) / 8.0
    regions = np.concatenate([c - h, c + h], 1)
    o64, l64 = s64.overlap(regions)
    o32, l32 = s32.overlap(regions.astype(np.float32))
    assert l64.size > 200 and np.array_equal(o64, o32) and np.array_equal(l64, l32)


def _balls_and_rays(dim, n, seed):
    rng = np.random.default_rng(seed)
    span = {2: 60.0, 3: 30.0, 5: 9.0}[dim]
    c = rng.uniform(-span, span, (n, dim))
    r = rng.uniform(0.3, 3.0, (n, 1))
    o = rng.uniform(-span - 5.0, span + 5.0, (2 * n, dim))
    d = rng.normal(size=o.shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([c, r], 1), np.concatenate([o, d], 1)


@pytest.mark.parametrize("dim", [2, 3, 5])
def test_consumer64_ball_equals_the_builtin_sphere(tmp_path, dim):
    """The oracle's f64 consumer kind running the built-in ball's arithmetic
    (BALL64, built by g++) equals its SPHERE kind bit for bit: nearest hits in
    the reference's walk and the naive scan, guarded and not, and all hits.  A
    record offset, a dropped range check or a guard on the wrong box fails it."""
    from consumer_narrow import BALL64, compile_cpu64
    balls, rays = _balls_and_rays(dim, 3000, 40 + dim)
    boxes = np.concatenate([balls[:, :dim] - balls[:, dim:], balls[:, :dim] + balls[:, dim:]], 1)
    _, addr = compile_cpu64(BALL64, dim, dim + 1, str(tmp_path))
    for guarded in (False, True):
        for naive in (False, True):
            cons = oracle.Consumer64(dim, boxes, balls, addr, guarded=guarded, naive=naive)
            ref = oracle.Scene64(dim, oracle.SPHERE, balls, guarded=guarded, naive=naive)
            a, b = cons.raycast_nearest(rays, 200.0), ref.raycast_nearest(rays, 200.0)
            assert (b["leaf"] >= 0).sum() > 500
            assert np.array_equal(a["leaf"], b["leaf"])
            assert np.array_equal(a["t"].view(np.uint64), b["t"].view(np.uint64))
            ca, cb = cons.raycast_all(rays, 200.0), ref.raycast_all(rays, 200.0)
            for x, y in zip(ca, cb):
                assert np.array_equal(np.asarray(x).view(np.uint64), np.asarray(y).view(np.uint64))


def test_consumer64_hits_outside_their_boxes(tmp_path):
    """Balls whose world boxes are shrunk to 0.6 of the radius (the reference
    asks only that the ray meet the leaf box, geometry.rs:37-45): the boxes'
    reject at the shrinking limit (the tree's nodes, the naive scan's cheap
    reject, naive.rs:34-62) then depends on the order the leaves are met, so
    the reference's walk and the naive scan answer differently on many rays,
    each answer a true hit of its ball.  With the guard both keep only hits
    inside their boxes and agree exactly."""
    from consumer_narrow import BALL64, compile_cpu64
    dim = 3
    balls, rays = _balls_and_rays(dim, 3000, 77)
    c, r = balls[:, :dim], balls[:, dim:]
    boxes = np.concatenate([c - 0.6 * r, c + 0.6 * r], 1)
    _, addr = compile_cpu64(BALL64, dim, dim + 1, str(tmp_path))
    walk = oracle.Consumer64(dim, boxes, balls, addr).raycast_nearest(rays, 200.0)
    naive = oracle.Consumer64(dim, boxes, balls, addr, naive=True).raycast_nearest(rays, 200.0)
    assert (walk["leaf"] != naive["leaf"]).sum() > 50
    for res in (walk, naive):
        hit = res["leaf"] >= 0
        assert hit.sum() > 1000
        p = rays[hit, :dim] + rays[hit, dim:] * res["t"][hit, None]
        dist = np.linalg.norm(p - c[res["leaf"][hit]], axis=1)
        assert np.allclose(dist, r[res["leaf"][hit], 0], rtol=1e-9, atol=1e-9)
    gw = oracle.Consumer64(dim, boxes, balls, addr, guarded=True).raycast_nearest(rays, 200.0)
    gn = oracle.Consumer64(dim, boxes, balls, addr, guarded=True, naive=True).raycast_nearest(rays, 200.0)
    assert np.array_equal(gw["leaf"], gn["leaf"])
    assert np.array_equal(gw["t"].view(np.uint64), gn["t"].view(np.uint64))


# Refit under the `f64` feature (Bvh::refit, bvh.rs:45-48; Tlas::refit and
# needs_rebuild, tlas.rs:49-60; tree::refit, tree.rs:54-75).  Pins: a refit to
# the same primitives is the build; a dyadic translation of every primitive
# (exact in double) refits to the tree the build over the moved primitives
# makes, so every answer equals the build's with the rays moved alike and the
# surface sum is unchanged; under random motion the refit tree answers what
# the naive scans over the moved primitives answer.  A refit that left any
# level of bounds stale loses the moved hits and fails the first two.

def _dyadic_prims(rng, n, dim, kind, span=2048):
    c = rng.integers(-span, span, (n, dim)) / 64.0
    if kind == oracle.BOX:
        return np.concatenate([c, c + rng.integers(8, 128, (n, dim)) / 64.0], 1)
    return np.concatenate([c, rng.integers(16, 96, (n, 1)) / 64.0], 1)


def _moved(prims, dim, kind, v):
    p = prims.copy()
    p[:, :dim] += v
    if kind == oracle.BOX:
        p[:, dim:2 * dim] += v
    return p


@pytest.mark.parametrize("dim,kind", [(2, oracle.BOX), (3, oracle.SPHERE), (6, oracle.BOX), (4, oracle.SPHERE)])
def test_bvh_refit_pins(dim, kind):
    rng = np.random.default_rng(910 + dim + kind)
    n = 3000
    span = {2: 2048, 3: 2048, 4: 512, 6: 128}[dim]   # denser in more dimensions, so rays still hit
    p0 = _dyadic_prims(rng, n, dim, kind, span)
    d = rng.normal(size=(2500, dim))
    d /= np.linalg.norm(d, axis=1, keepdims=True)   # unit directions (Ray::new), as the ball test assumes
    rays = np.concatenate([rng.integers(-span - span // 8, span + span // 8, (2500, dim)) / 64.0, d], 1)
    s0 = oracle.Scene64(dim, kind, p0)
    built = s0.bvh_raycast_nearest(rays, 200.0)
    q0 = s0.tree_stats()["quality"]
    same = s0.bvh_raycast_nearest_refit(p0, rays, 200.0)
    assert np.array_equal(same["leaf"], built["leaf"]) and np.array_equal(same["t"], built["t"])
    assert same["quality"] == q0
    assert (built["leaf"] >= 0).sum() > 400
    # dyadic translation: the refit tree is the moved build
    v = np.array([17.25, -3.5, 6.125, -0.75, 2.0, 9.5][:dim])
    p1 = _moved(p0, dim, kind, v)
    r1 = rays.copy(); r1[:, :dim] += v
    s1 = oracle.Scene64(dim, kind, p1)
    ref = s1.bvh_raycast_nearest_refit(p0, r1, 200.0)
    direct = s1.bvh_raycast_nearest(r1, 200.0)
    assert np.array_equal(ref["leaf"], built["leaf"]) and np.array_equal(ref["t"], built["t"])
    assert np.array_equal(direct["leaf"], built["leaf"]) and np.array_equal(direct["t"], built["t"])
    assert ref["quality"] == q0 == s1.tree_stats()["quality"]
    # random motion: the refit tree answers the naive scan over the moved primitives
    p2 = _moved(p0, dim, kind, rng.uniform(-4.0, 4.0, (n, dim)))
    s2 = oracle.Scene64(dim, kind, p2)
    got = s2.bvh_raycast_nearest_refit(p0, rays, 200.0)
    want = s2.raycast_nearest(rays, 200.0)
    assert np.array_equal(got["leaf"], want["leaf"]) and np.array_equal(got["t"], want["t"])
    assert got["quality"] > q0   # the moved boxes spread the build's nodes


def _signed_permutations(rng, n, dim):
    rot = np.zeros((n, dim, dim))
    for i in range(n):
        rot[i, np.arange(dim), rng.permutation(dim)] = rng.choice([-1.0, 1.0], dim)
    return rot


@pytest.mark.parametrize("dim,kind", [(3, 0), (2, 0), (3, 1)])
def test_tlas_refit_pins(dim, kind):
    rng = np.random.default_rng(930 + dim + kind)
    n = 400
    rot = _signed_permutations(rng, n, dim)
    t0 = rng.integers(-1024, 1024, (n, dim)) / 32.0
    half = rng.integers(8, 64, (n, dim)) / 32.0 if kind == 0 else rng.integers(16, 64, n) / 32.0
    d = rng.normal(size=(2500, dim))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([rng.integers(-1200, 1200, (2500, dim)) / 32.0, d], 1)
    built = oracle.Instances64(dim, kind, rot, t0, half)
    q0 = built.quality()
    same = oracle.Instances64(dim, kind, rot, t0, half, built_from=(rot, t0, half))
    a, b = built.raycast_nearest(rays, 300.0), same.raycast_nearest(rays, 300.0)
    assert np.array_equal(a["leaf"], b["leaf"]) and np.array_equal(a["t"], b["t"]) and same.quality() == q0
    assert (a["leaf"] >= 0).sum() > 300
    # dyadic translation of every instance: the refit tree is the moved build
    v = np.array([40.5, -12.25, 3.0][:dim])
    r1 = rays.copy(); r1[:, :dim] += v
    moved = oracle.Instances64(dim, kind, rot, t0 + v, half, built_from=(rot, t0, half))
    direct = oracle.Instances64(dim, kind, rot, t0 + v, half)
    m, d = moved.raycast_nearest(r1, 300.0), direct.raycast_nearest(r1, 300.0)
    assert np.array_equal(m["leaf"], a["leaf"]) and np.array_equal(m["t"], a["t"])
    assert np.array_equal(d["leaf"], a["leaf"]) and np.array_equal(d["t"], a["t"])
    assert moved.quality() == q0 == direct.quality()
    assert not moved.needs_rebuild(1.0, q0)
    # random motion and new rotations: all hits (one crossing per instance, in
    # any tree order) and guarded nearest hits equal the naive scans
    t2 = t0 + rng.uniform(-6.0, 6.0, (n, dim))
    rot2 = _plane_rotations(rng, n, dim)
    moved = oracle.Instances64(dim, kind, rot2, t2, half, built_from=(rot, t0, half))
    naive = oracle.Instances64(dim, kind, rot2, t2, half, naive=True)
    oa, la, ta = moved.raycast_all(rays, 300.0)
    ob, lb, tb = naive.raycast_all(rays, 300.0)
    assert np.array_equal(oa, ob) and np.array_equal(la, lb) and np.array_equal(ta, tb) and la.size > 300
    g = oracle.Instances64(dim, kind, rot2, t2, half, guarded=True, built_from=(rot, t0, half))
    gn = oracle.Instances64(dim, kind, rot2, t2, half, guarded=True, naive=True)
    x, y = g.raycast_nearest(rays, 300.0), gn.raycast_nearest(rays, 300.0)
    assert np.array_equal(x["leaf"], y["leaf"]) and np.array_equal(x["t"], y["t"])
    assert moved.quality() > q0 and moved.needs_rebuild(1.0, q0)


def test_consumer64_shell_crossings(tmp_path):
    """Bvh::raycast_crossings (bvh.rs:264-328) over a consumer in double whose
    test_ray_crossings reports both roots (the shells of crossings.rs:38-111):
    closed forms (a shell of radius 1 at x = 5 crossed at 4 and 6 from the
    origin, at 1 only from its centre, not at all by a ray that passes 2 above
    it, at 4 only when max_toi is 5); the BVH walk's lists equal the naive
    scan's; every ray's first crossing is its nearest hit, and the all-hits
    list (one test_ray hit per shell) is a subset of the crossings."""
    from consumer_narrow import SHELL64, compile_cpu64, crossings64_addr
    lib, addr = compile_cpu64(SHELL64, 3, 4, str(tmp_path))
    xa = crossings64_addr(lib)
    assert xa
    one = np.array([[5.0, 0.0, 0.0, 1.0]])
    box = np.array([[4.0, -1.0, -1.0, 6.0, 1.0, 1.0]])
    c = oracle.Consumer64(3, box, one, addr)
    rays = np.array([[0.0, 0, 0, 1, 0, 0], [5.0, 0, 0, 1, 0, 0], [0.0, 2, 0, 1, 0, 0]])
    off, leaf, t = c.raycast_crossings(rays, 100.0, xa)
    assert off.tolist() == [0, 2, 3, 3] and leaf.tolist() == [0, 0, 0] and t.tolist() == [4.0, 6.0, 1.0]
    off, leaf, t = c.raycast_crossings(rays[:1], 5.0, xa)
    assert off.tolist() == [0, 1] and t.tolist() == [4.0]
    rng = np.random.default_rng(0x5E11)
    n = 2000
    ctr = rng.uniform(-40.0, 40.0, (n, 3)); rad = rng.uniform(0.5, 3.0, (n, 1))
    prims = np.concatenate([ctr, rad], 1)
    boxes = np.concatenate([ctr - rad, ctr + rad], 1)
    o = rng.uniform(-45.0, 45.0, (3000, 3)); d = rng.normal(size=(3000, 3))
    q = np.concatenate([o, d / np.linalg.norm(d, axis=1, keepdims=True)], 1)
    bvh = oracle.Consumer64(3, boxes, prims, addr)
    naive = oracle.Consumer64(3, boxes, prims, addr, naive=True)
    ob, lb, tb = bvh.raycast_crossings(q, 150.0, xa)
    on, ln, tn = naive.raycast_crossings(q, 150.0, xa)
    assert np.array_equal(ob, on) and np.array_equal(lb, ln) and np.array_equal(tb, tn) and lb.size > 3000
    near = naive.raycast_nearest(q, 150.0)
    oa, la, ta = naive.raycast_all(q, 150.0)
    for i in range(len(q)):
        xs = list(zip(tb[ob[i]:ob[i + 1]].tolist(), lb[ob[i]:ob[i + 1]].tolist()))
        if near["leaf"][i] >= 0:
            assert xs[0] == (near["t"][i], near["leaf"][i])
        else:
            assert not xs
        assert set(zip(ta[oa[i]:oa[i + 1]].tolist(), la[oa[i]:oa[i + 1]].tolist())) <= set(xs)
