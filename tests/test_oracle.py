"""Pins the CPU oracle: the reference's own tests restated (same SplitMix64
draws, same providers, same assertions), closed forms, and float64 textbook
reductions that a dropped term or a wrong sign would fail.

Reference tests restated here (crates/spatial-query/tests/):
  differential.rs   BVH == naive bit-exact at D = 2, 3, 4; edge cases; ties
  conformance.rs    repeatability; parallel == sequential for every thread count
  two_level.rs      isometry / transform_aabb closed forms; Tlas == naive;
                    refit == rebuild; needs_rebuild; Bvh refit == rebuild
  shape_overlap.rs  swept ball toi 8.25; overlap [0] and [0, 1]; BVH == naive
  crossings.rs      shell entry/exit 8 and 12; BVH == naive
  result_identity.rs filter complement
  stats.rs          tree stats of one-leaf / empty trees; exact counters
"""
import math

import numpy as np
import pytest

import oracle
from sq_amd.workloads import Stream, normalize_or_zero

F32 = np.float32


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def ray_new(o, d):
    return np.concatenate([np.asarray(o, F32), normalize_or_zero(np.asarray(d, F32)[None])[0]])[None, :]


# ---------------------------------------------------------------------------
# differential.rs
# ---------------------------------------------------------------------------

def _parity(prims, dim, ray, max_toi, kind=oracle.SPHERE_PLAIN):
    s = oracle.Scene(dim, kind, prims)
    a = s.raycast_nearest(ray, max_toi)
    b = s.raycast_nearest(ray, max_toi, naive=True)
    assert a["leaf"][0] == b["leaf"][0]
    assert bits(a["t"])[0] == bits(b["t"])[0]
    oa, la, ta = s.raycast_all(ray, max_toi)
    ob, lb, tb = s.raycast_all(ray, max_toi, naive=True)
    assert np.array_equal(la, lb) and np.array_equal(bits(ta), bits(tb))
    return a, la


@pytest.mark.parametrize("dim,seed,n,span", [(3, 0xABCD_1234, 64, 60.0), (2, 0x2222_2222, 48, 60.0), (4, 0x4444_4444, 48, 40.0)])
def test_bvh_matches_naive(dim, seed, n, span):
    """differential.rs:120-162: 200 random ball scenes per dimension."""
    s = Stream(seed)
    hits = 0
    for _ in range(200):
        c = s.range(-50.0, 50.0, n * dim).reshape(n, dim)
        r = s.range(0.5, 4.0, n)
        o = s.range(-span, span, dim)
        d = s.range(-1.0, 1.0, dim)
        ray = ray_new(o, d)
        if not ray[0, dim:].any():
            continue
        prims = np.concatenate([c, r[:, None]], 1)
        a, _ = _parity(prims, dim, ray, 500.0)
        hits += a["leaf"][0] >= 0
    assert hits > 5 or dim == 4   # the reference D=4 draws are sparse: every ray misses


def test_edge_cases():
    """differential.rs:164-194."""
    ray = ray_new([0, 0, 0], [1, 0, 0])
    empty = oracle.Scene(3, oracle.SPHERE_PLAIN, np.zeros((0, 4), F32))
    assert empty.raycast_nearest(ray, 100.0)["leaf"][0] == -1
    assert len(empty.raycast_all(ray, 100.0)[1]) == 0
    one = oracle.Scene(3, oracle.SPHERE_PLAIN, np.array([[5, 0, 0, 1]], F32))
    h = one.raycast_nearest(ray, 100.0)
    assert h["leaf"][0] == 0 and abs(h["t"][0] - 4.0) < 1e-4
    assert one.raycast_nearest(ray_new([0, 0, 0], [-1, 0, 0]), 100.0)["leaf"][0] == -1
    assert one.raycast_nearest(ray, 3.0)["leaf"][0] == -1


def test_ties_sorted_by_leaf():
    """differential.rs:196-211: equal toi -> leaf order."""
    s = oracle.Scene(3, oracle.SPHERE_PLAIN, np.array([[5, 0, 0, 1], [5, 0, 0, 1]], F32))
    off, leaf, t = s.raycast_all(ray_new([0, 0, 0], [1, 0, 0]), 100.0)
    assert leaf.tolist() == [0, 1] and t[0] <= t[1]
    assert s.raycast_nearest(ray_new([0, 0, 0], [1, 0, 0]), 100.0)["leaf"][0] == 0


def test_doc_example():
    """lib.rs:21-55: balls at 5 and 10, unit radius; hit id 0 at toi 4."""
    s = oracle.Scene(3, oracle.SPHERE_PLAIN_NOEXIT, np.array([[5, 0, 0, 1], [10, 0, 0, 1]], F32))
    h = s.raycast_nearest(ray_new([0, 0, 0], [1, 0, 0]), 100.0)
    assert h["leaf"][0] == 0 and abs(h["t"][0] - 4.0) < 1e-4


# ---------------------------------------------------------------------------
# conformance.rs
# ---------------------------------------------------------------------------

def _scatter(n, seed):
    s = Stream(seed)
    out = np.empty((n, 4), F32)
    for i in range(n):
        out[i, :3] = s.range(-20.0, 20.0, 3)
        out[i, 3] = s.range(0.3, 1.4, 1)[0]
    return out


def _fan(n, seed):
    s = Stream(seed)
    rays = []
    for _ in range(n):
        o = s.range(-30.0, 30.0, 3)
        tgt = s.range(-8.0, 8.0, 3)
        rays.append(ray_new(o, tgt - o)[0])
    return np.array(rays, F32)


def test_repeatable_and_parallel_equals_sequential():
    """conformance.rs:106-181."""
    balls = oracle.Scene(3, oracle.SPHERE_PLAIN_NOEXIT, _scatter(800, 0x1234_5678_9ABC_DEF0))
    rays = _fan(500, 0xFACE_B00C_1357_9BDF)
    seq = balls.raycast_nearest(rays, 200.0)
    assert (seq["leaf"] >= 0).any() and (seq["leaf"] < 0).any()
    for _ in range(20):
        again = balls.raycast_nearest(rays, 200.0)
        assert np.array_equal(again["leaf"], seq["leaf"]) and np.array_equal(bits(again["t"]), bits(seq["t"]))
    for threads in [1, 2, 3, 4, 7, 8, 16, 999]:
        par = balls.raycast_nearest(rays, 200.0, threads=threads)
        for k in ("leaf", "sub"):
            assert np.array_equal(par[k], seq[k])
        for k in ("t", "point", "normal"):
            assert np.array_equal(bits(par[k]), bits(seq[k]))
    # single rays equal the batch
    for i in range(0, 500, 37):
        one = balls.raycast_nearest(rays[i:i + 1], 200.0)
        assert one["leaf"][0] == seq["leaf"][i] and bits(one["t"])[0] == bits(seq["t"])[i]
    empty = balls.raycast_nearest(np.zeros((0, 6), F32), 200.0, threads=8)
    assert len(empty["leaf"]) == 0


# ---------------------------------------------------------------------------
# two_level.rs
# ---------------------------------------------------------------------------

def plane_rotation(dim, a, b, angle):
    """isometry.rs:129-140 (sin/cos in f32 of the f32 angle)."""
    r = np.eye(dim, dtype=F32)
    sn, cs = F32(math.sin(float(angle))), F32(math.cos(float(angle)))
    r[a, a] = cs; r[a, b] = -sn; r[b, a] = sn; r[b, b] = cs
    return r


def test_isometry_closed_forms():
    """two_level.rs:180-210."""
    rot = plane_rotation(3, 0, 1, math.pi / 2)
    t0 = np.zeros(3, F32)
    r = oracle.iso_apply(rot, t0, np.array([1, 0, 0], F32), 1)
    assert abs(r[0]) < 1e-6 and abs(r[1] - 1) < 1e-6 and abs(r[2]) < 1e-6
    back = oracle.iso_apply(rot, t0, r, 3)
    assert abs(back[0] - 1) < 1e-6 and abs(back[1]) < 1e-6
    rot = plane_rotation(3, 0, 1, math.pi / 4)
    w = oracle.iso_transform_aabb(rot, np.array([10, 0, 0], F32), np.array([-.5, -.5, -.5, .5, .5, .5], F32))
    ext = w[3:] - w[:3]
    diag = 2.0 * (1 / math.sqrt(2))
    assert abs(ext[0] - diag) < 1e-5 and abs(ext[1] - diag) < 1e-5 and abs(ext[2] - 1) < 1e-5
    assert abs((w[0] + w[3]) * 0.5 - 10) < 1e-5
    # inverse_transform_point(transform_point(p)) == p (orthonormal R) within rounding
    p = np.array([1.5, -2.0, 3.25], F32)
    tr = np.array([4, 5, 6], F32)
    q = oracle.iso_apply(rot, tr, oracle.iso_apply(rot, tr, p, 0), 2)
    assert np.allclose(q, p, atol=1e-5)


def _random_boxes(s, dim, n):
    """two_level.rs:115-131."""
    rot = np.empty((n, dim, dim), F32); tr = np.empty((n, dim), F32); half = np.empty((n, dim), F32)
    for i in range(n):
        tr[i] = s.range(-40.0, 40.0, dim)
        a = int(s.u64(1)[0] % np.uint64(dim))
        b = int(s.u64(1)[0] % np.uint64(dim))
        if b == a:
            b = (b + 1) % dim
        rot[i] = plane_rotation(dim, a, b, s.range(0.0, 2 * math.pi, 1)[0])
        half[i] = np.maximum(s.range(0.5, 2.5, dim), F32(0.5))
    return rot, tr, half


def test_end_to_end_box_instance():
    """two_level.rs:212-224: unit box at x = 5 -> toi 4, point x 4, normal -x."""
    sc = oracle.InstanceScene(3, oracle.INST_BOX, np.eye(3, dtype=F32)[None], np.array([[5, 0, 0]], F32),
                              halves=np.ones((1, 3), F32))
    h = sc.raycast_nearest(ray_new([0, 0, 0], [1, 0, 0]), 100.0)
    assert h["leaf"][0] == 0 and abs(h["t"][0] - 4) < 1e-4
    assert abs(h["point"][0, 0] - 4) < 1e-4 and abs(h["normal"][0, 0] + 1) < 1e-4


@pytest.mark.parametrize("dim,seed,n", [(3, 0x7ABC_1234, 64), (2, 0x2020_2020, 48)])
def test_tlas_matches_naive(dim, seed, n):
    """two_level.rs:263-292."""
    s = Stream(seed)
    hits = 0
    for _ in range(200):
        rot, tr, half = _random_boxes(s, dim, n)
        ray = ray_new(s.range(-60.0, 60.0, dim), s.range(-1.0, 1.0, dim))
        sc = oracle.InstanceScene(dim, oracle.INST_BOX, rot, tr, halves=half)
        a = sc.raycast_nearest(ray, 500.0)
        b = sc.raycast_nearest(ray, 500.0, naive=True)
        assert a["leaf"][0] == b["leaf"][0] and bits(a["t"])[0] == bits(b["t"])[0]
        off, leaf, t, _ = sc.raycast_all(ray, 500.0)
        order = np.lexsort((leaf, t))   # naive_all sorts by (toi, leaf)
        assert np.array_equal(order, np.arange(len(leaf)))
        hits += a["leaf"][0] >= 0
    assert hits > 0


def test_tlas_refit_matches_rebuild_and_naive():
    """two_level.rs:294-342."""
    s = Stream(0xF00D_5EED)
    rot, tr, half = _random_boxes(s, 3, 200)
    sc = oracle.InstanceScene(3, oracle.INST_BOX, rot, tr, halves=half)
    for frame in range(10):
        for i in range(200):
            tr[i] = tr[i] + s.range(-1.5, 1.5, 3)
            if i % 7 == 0:
                half[i] = half[i] * F32(1.05)
        sc.set_transforms(rot, tr, half)
        sc.refit()
        fresh = oracle.InstanceScene(3, oracle.INST_BOX, rot, tr, halves=half)
        rays = []
        for _ in range(40):
            rays.append(ray_new(s.range(-60.0, 60.0, 3), s.range(-1.0, 1.0, 3))[0])
        rays = np.array(rays, F32)
        a = sc.raycast_nearest(rays, 500.0)
        b = fresh.raycast_nearest(rays, 500.0)
        c = sc.raycast_nearest(rays, 500.0, naive=True)
        for x in (b, c):
            assert np.array_equal(a["leaf"], x["leaf"]) and np.array_equal(bits(a["t"]), bits(x["t"]))


def test_needs_rebuild_after_spreading():
    """two_level.rs:344-372."""
    s = Stream(0xCAFE_1111)
    n = 128
    tr = s.range(-2.0, 2.0, n * 3).reshape(n, 3)
    rot = np.tile(np.eye(3, dtype=F32), (n, 1, 1))
    sc = oracle.InstanceScene(3, oracle.INST_BOX, rot, tr, halves=np.full((n, 3), 0.5, F32))
    assert not sc.needs_rebuild(2.0)
    sc.set_transforms(rot, s.range(-500.0, 500.0, n * 3).reshape(n, 3))
    sc.refit()
    assert sc.needs_rebuild(2.0)


def test_bvh_refit_matches_rebuild():
    """two_level.rs:405-432 (unit balls, 150 centres, 8 frames)."""
    s = Stream(0x1357_9BDF)
    c = s.range(-30.0, 30.0, 150 * 3).reshape(150, 3)
    prims = np.concatenate([c, np.ones((150, 1), F32)], 1)
    sc = oracle.Scene(3, oracle.SPHERE_PLAIN_NOEXIT, prims)
    for _ in range(8):
        for i in range(150):
            prims[i, :3] = prims[i, :3] + s.range(-2.0, 2.0, 3)
        sc.set_prims(prims)
        sc.refit()
        fresh = oracle.Scene(3, oracle.SPHERE_PLAIN_NOEXIT, prims)
        rays = np.array([ray_new(s.range(-40.0, 40.0, 3), s.range(-1.0, 1.0, 3))[0] for _ in range(40)], F32)
        a = sc.raycast_nearest(rays, 300.0)
        b = fresh.raycast_nearest(rays, 300.0)
        assert np.array_equal(a["leaf"], b["leaf"]) and np.array_equal(bits(a["t"]), bits(b["t"]))
        # and the refit bounds equal the rebuilt tree's where the topology agrees: the root
        assert np.array_equal(sc.export()[0][0], fresh.export()[0][0])


# ---------------------------------------------------------------------------
# shape_overlap.rs, crossings.rs, result_identity.rs, stats.rs
# ---------------------------------------------------------------------------

def ball_cast(o, d, probe=0.75):
    d = normalize_or_zero(np.asarray(d, F32)[None])[0]
    return np.concatenate([np.asarray(o, F32), d, np.full(3, -probe, F32), np.full(3, probe, F32)])[None]


def test_swept_ball_toi():
    """shape_overlap.rs:184-230: probe 0.75 vs ball r 1 at x 10 -> toi 8.25, normal -x."""
    sc = oracle.Scene(3, oracle.SPHERE_PLAIN, np.array([[10, 0, 0, 1]], F32))
    h = sc.shapecast_nearest(ball_cast([0, 0, 0], [1, 0, 0]), 100.0)
    assert h["leaf"][0] == 0 and abs(h["t"][0] - 8.25) < 1e-4 and abs(h["normal"][0, 0] + 1) < 1e-4
    # fine-step march (the reference cross-check)
    t = 0.0
    while math.dist((t, 0, 0), (10, 0, 0)) > 1.75:
        t += 0.001
    assert abs(h["t"][0] - t) < 0.002


def test_overlap_known_answers():
    """shape_overlap.rs:236-262."""
    sc = oracle.Scene(3, oracle.SPHERE_PLAIN, np.array([[0, 0, 0, 1], [5, 0, 0, 1], [100, 0, 0, 1]], F32))
    off, leaf = sc.overlap(np.array([[-1.5, -1.5, -1.5, 1.5, 1.5, 1.5]], F32))
    assert leaf.tolist() == [0]
    off, leaf = sc.overlap(np.array([[-1.5, -1.5, -1.5, 4.0, 1.5, 1.5]], F32))
    assert leaf.tolist() == [0, 1]


def test_shapecast_and_overlap_match_naive():
    """shape_overlap.rs:153-182, 264-285."""
    s = Stream(0x5EED_C0DE)
    for _ in range(100):
        c = s.range(-30.0, 30.0, 64 * 3).reshape(64, 3)
        r = s.range(0.5, 3.0, 64)
        prims = np.concatenate([c, r[:, None]], 1)
        sc = oracle.Scene(3, oracle.SPHERE_PLAIN, prims)
        cast = ball_cast(s.range(-50.0, 50.0, 3), s.range(-1.0, 1.0, 3))
        a = sc.shapecast_nearest(cast, 500.0); b = sc.shapecast_nearest(cast, 500.0, naive=True)
        assert a["leaf"][0] == b["leaf"][0] and bits(a["t"])[0] == bits(b["t"])[0]
        oa, la, ta = sc.shapecast_all(cast, 500.0); ob, lb, tb = sc.shapecast_all(cast, 500.0, naive=True)
        assert np.array_equal(la, lb) and np.array_equal(bits(ta), bits(tb))
        ctr = s.range(-30.0, 30.0, 3); h = s.range(1.0, 6.0, 3)
        reg = np.concatenate([ctr - h, ctr + h])[None]
        assert np.array_equal(sc.overlap(reg)[1], sc.overlap(reg, naive=True)[1])


def ball_instances(s, n, dim=3, lo=-30.0, hi=30.0, rlo=0.5, rhi=2.0):
    """BallInstances (shape_overlap.rs:288-305): translations only, radii U[0.5, 2]."""
    rot = np.tile(np.eye(dim, dtype=F32), (n, 1, 1))
    trans = s.range(lo, hi, n * dim).reshape(n, dim).astype(F32)
    radii = s.range(rlo, rhi, n).astype(F32)
    return rot, trans, radii


def test_tlas_shapecast_matches_naive_via_ball_bvh():
    """shape_overlap.rs:347-382: a TLAS of ball instances answers shape casts
    like the naive scan over the same balls as a single-level provider."""
    s = Stream(0x1234_ABCD)
    for _ in range(150):
        rot, trans, radii = ball_instances(s, 48)
        inst = oracle.InstanceScene(3, oracle.INST_BALL, rot, trans, halves=radii)
        balls = oracle.Scene(3, oracle.SPHERE, np.concatenate([trans, radii[:, None]], 1))
        cast = ball_cast(s.range(-50.0, 50.0, 3), s.range(-1.0, 1.0, 3))
        want = balls.shapecast_nearest(cast, 500.0, naive=True)
        got = inst.shapecast_nearest(cast, 500.0)
        assert got["leaf"][0] == want["leaf"][0] and bits(got["t"])[0] == bits(want["t"])[0]
        _, lw, tw = balls.shapecast_all(cast, 500.0, naive=True)
        _, lg, tg = inst.shapecast_all(cast, 500.0)
        assert np.array_equal(lg, lw) and np.array_equal(bits(tg), bits(tw))


def test_tlas_overlap_matches_naive_via_ball_bvh():
    """shape_overlap.rs:384-411: the instances' exact test_overlap."""
    s = Stream(0x9999_5555)
    hits = 0
    for _ in range(150):
        rot, trans, radii = ball_instances(s, 64)
        inst = oracle.InstanceScene(3, oracle.INST_BALL, rot, trans, halves=radii)
        balls = oracle.Scene(3, oracle.SPHERE, np.concatenate([trans, radii[:, None]], 1))
        c = s.range(-30.0, 30.0, 3); h = s.range(1.0, 6.0, 3)
        reg = np.concatenate([c - h, c + h])[None].astype(F32)
        want = balls.overlap(reg, naive=True)[1]
        got = inst.overlap(reg)[1]
        assert np.array_equal(got, want)
        hits += len(got)
    assert hits > 20


def test_tlas_ball_rays_match_single_level_balls():
    """Rays through ball instances (ray_ball on the local ray, shape_overlap.rs:
    323-325) give the single-level balls' hits, and a box instance kind
    answers no shape cast (the trait default, instance.rs:72-87)."""
    s = Stream(0xBA11)
    rot, trans, radii = ball_instances(s, 500, lo=-60.0, hi=60.0)
    inst = oracle.InstanceScene(3, oracle.INST_BALL, rot, trans, halves=radii)
    balls = oracle.Scene(3, oracle.SPHERE, np.concatenate([trans, radii[:, None]], 1))
    d = normalize_or_zero(s.range(-1.0, 1.0, 2000 * 3).reshape(-1, 3).astype(F32))
    rays = np.concatenate([s.range(-70.0, 70.0, 2000 * 3).reshape(-1, 3).astype(F32), d], 1)
    a, b = inst.raycast_nearest(rays, 300.0), balls.raycast_nearest(rays, 300.0)
    assert (a["leaf"] >= 0).sum() > 100
    assert np.array_equal(a["leaf"], b["leaf"]) and np.array_equal(bits(a["t"]), bits(b["t"]))
    boxes = oracle.InstanceScene(3, oracle.INST_BOX_GUARDED, rot, trans, halves=np.tile(radii[:, None], (1, 3)))
    casts = np.concatenate([rays[:50, :6], np.full((50, 3), -0.5, F32), np.full((50, 3), 0.5, F32)], 1)
    assert (boxes.shapecast_nearest(casts, 300.0)["leaf"] == -1).all()
    assert boxes.shapecast_all(casts, 300.0)[0][-1] == 0


def test_crossings_entry_exit():
    """crossings.rs:113-135: a shell of radius 2 at x 10 is crossed at 8 and 12."""
    sc = oracle.Scene(3, oracle.SHELL, np.array([[10, 0, 0, 2]], F32))
    off, leaf, t = sc.raycast_all(ray_new([0, 0, 0], [1, 0, 0]), 100.0, crossings=True)
    assert leaf.tolist() == [0, 0] and abs(t[0] - 8) < 1e-4 and abs(t[1] - 12) < 1e-4
    # default crossings (test_ray_crossings not overridden) equal all-hits
    b = oracle.Scene(3, oracle.SPHERE_PLAIN, np.array([[10, 0, 0, 2]], F32))
    assert np.array_equal(b.raycast_all(ray_new([0, 0, 0], [1, 0, 0]), 100.0, crossings=True)[1],
                          b.raycast_all(ray_new([0, 0, 0], [1, 0, 0]), 100.0)[1])


@pytest.mark.parametrize("dim,seed", [(3, 0xC055_1234), (2, 0x2C05_5222)])
def test_crossings_match_naive(dim, seed):
    """crossings.rs:137-200."""
    s = Stream(seed)
    for _ in range(100):
        c = s.range(-30.0, 30.0, 48 * dim).reshape(48, dim)
        r = s.range(1.0, 4.0, 48)
        sc = oracle.Scene(dim, oracle.SHELL, np.concatenate([c, r[:, None]], 1))
        ray = ray_new(s.range(-40.0, 40.0, dim), s.range(-1.0, 1.0, dim))
        oa, la, ta = sc.raycast_all(ray, 500.0, crossings=True)
        ob, lb, tb = sc.raycast_all(ray, 500.0, crossings=True, naive=True)
        assert np.array_equal(la, lb) and np.array_equal(bits(ta), bits(tb))


def test_filter_selects_complement():
    """result_identity.rs:123-163: even leaves on layer bit 0, odd on bit 1."""
    n = 12
    prims = np.array([[2.0 * (i + 1), 0, 0, 0.6] for i in range(n)], F32)
    layers = np.array([1 << (i % 2) for i in range(n)], np.uint64)
    sc = oracle.Scene(3, oracle.SPHERE_PLAIN, prims, layers=layers)
    ray = ray_new([0, 0, 0], [1, 0, 0])
    even = sc.raycast_all(ray, 100.0, mask=0b01)[1].tolist()
    odd = sc.raycast_all(ray, 100.0, mask=0b10)[1].tolist()
    both = sc.raycast_all(ray, 100.0)[1].tolist()
    assert even == list(range(0, n, 2)) and odd == list(range(1, n, 2)) and both == list(range(n))
    assert sc.raycast_nearest(ray, 100.0, mask=0b10)["leaf"][0] == 1
    assert sc.raycast_nearest(ray, 100.0, exclude=0)["leaf"][0] == 1
    reg = np.array([[-100, -1, -1, 100, 1, 1]], F32)
    assert sc.overlap(reg, mask=0b01)[1].tolist() == list(range(0, n, 2))


def test_tree_stats_and_counters():
    """stats.rs:62-185 (balls of radius 0.5 in a line at x = 2(i+1))."""
    line = lambda n: np.array([[2.0 * (i + 1), 0, 0, 0.5] for i in range(n)], F32)
    s = oracle.Scene(3, oracle.SPHERE_PLAIN_NOEXIT, line(3)).tree_stats()
    assert (s["node_count"], s["leaf_count"], s["prim_count"], s["max_depth"]) == (1, 1, 3, 1) and s["bytes"] > 0
    s = oracle.Scene(3, oracle.SPHERE_PLAIN_NOEXIT, line(50)).tree_stats()
    assert s["prim_count"] == 50 and 1 <= s["leaf_count"] < s["node_count"] and 1 <= s["max_depth"] <= s["node_count"]
    s = oracle.Scene(3, oracle.SPHERE_PLAIN_NOEXIT, line(0)).tree_stats()
    assert (s["node_count"], s["leaf_count"], s["prim_count"], s["max_depth"]) == (0, 0, 0, 0)
    sc = oracle.Scene(3, oracle.SPHERE_PLAIN_NOEXIT, line(3))
    st = sc.raycast_nearest(ray_new([0, 0, 0], [1, 0, 0]), 100.0, stats=True)["stats"][0]
    # one node, one aabb test, three narrow tests, three hits (each closer... no: nearest keeps testing all)
    assert st.tolist()[:3] == [1, 1, 3]
    miss = sc.raycast_nearest(ray_new([0, 5, 0], [1, 0, 0]), 100.0, stats=True)["stats"][0]
    assert miss.tolist() == [1, 1, 0, 0]
    # all_hits_counters_are_exact (stats.rs:107-122): one node, one test, three narrow tests, three hits
    assert sc.raycast_all_stats(ray_new([0, 0, 0], [1, 0, 0]), 100.0)[0].tolist() == [1, 1, 3, 3]
    assert sc.raycast_all_stats(ray_new([0, 5, 0], [1, 0, 0]), 100.0)[0].tolist() == [1, 1, 0, 0]


def test_query_counters_on_a_larger_tree():
    """stats_path_matches_plain_path_and_repeats (stats.rs:165-185) and what
    the walk fixes: every popped node is tested (visits == AABB tests), both
    children of every passing interior node are popped (visits odd: 1 + 2k),
    a hit is a narrow test that hit, and the all-hits walk, which never prunes,
    tests every node the nearest walk tests and more."""
    rng = np.random.default_rng(11)
    c = rng.uniform(-30, 30, (400, 3)).astype(F32)
    balls = np.concatenate([c, np.full((400, 1), 1.0, F32)], 1)
    sc = oracle.Scene(3, oracle.SPHERE, balls)
    o = rng.uniform(-50, 50, (300, 3)).astype(F32)
    rays = np.concatenate([o, -o / np.linalg.norm(o, axis=1, keepdims=True)], 1).astype(F32)
    a = sc.raycast_nearest(rays, 200.0, stats=True)
    b = sc.raycast_nearest(rays, 200.0, stats=True)
    assert np.array_equal(a["stats"], b["stats"]) and np.array_equal(a["leaf"], sc.raycast_nearest(rays, 200.0)["leaf"])
    st = a["stats"].astype(np.int64)
    nn = sc.tree_stats()["node_count"]
    assert (st[:, 0] == st[:, 1]).all() and (st[:, 0] % 2 == 1).all()
    assert (st[:, 3] <= st[:, 2]).all() and (st[:, 3] >= (a["leaf"] >= 0)).all()
    assert (st[:, 0] <= st[:, 2] + nn).all() and st[:, 0].max() > 1
    al = sc.raycast_all_stats(rays, 200.0).astype(np.int64)
    off, _, _ = sc.raycast_all(rays, 200.0)
    assert np.array_equal(al[:, 3], np.diff(off.astype(np.int64)))
    assert (al[:, 0] >= st[:, 0]).all() and (al[:, 2] >= st[:, 2]).all() and (al[:, 0] % 2 == 1).all()


# ---------------------------------------------------------------------------
# narrow phases and the slab against closed forms / float64 reductions
# ---------------------------------------------------------------------------

def test_slab_closed_forms():
    box = np.array([1, 1, 1, 3, 3, 3], F32)
    assert oracle.ray_intersect(box, np.array([0, 2, 2, 1, 0, 0], F32), 100.0) == (1.0, 3.0)
    assert oracle.ray_intersect(box, np.array([0, 2, 2, 1, 0, 0], F32), 0.5) is None      # max_toi clips
    assert oracle.ray_intersect(box, np.array([2, 2, 2, 1, 0, 0], F32), 100.0) == (0.0, 1.0)   # inside: t_min 0
    assert oracle.ray_intersect(box, np.array([0, 5, 2, 1, 0, 0], F32), 100.0) is None      # parallel, outside
    assert oracle.ray_intersect(box, np.array([0, 3, 2, 1, 0, 0], F32), 100.0) == (1.0, 3.0)  # parallel, on face


def test_slab_matches_float64_intervals():
    rng = np.random.default_rng(3)
    for _ in range(2000):
        c = rng.uniform(-10, 10, 3); h = rng.uniform(0.1, 3, 3)
        box = np.concatenate([c - h, c + h]).astype(F32)
        o = rng.uniform(-20, 20, 3).astype(F32)
        d = normalize_or_zero(rng.standard_normal((1, 3)).astype(F32))[0]
        got = oracle.ray_intersect(box, np.concatenate([o, d]), 1000.0)
        b64, o64, d64 = box.astype(np.float64), o.astype(np.float64), d.astype(np.float64)
        t1 = (b64[:3] - o64) / d64; t2 = (b64[3:] - o64) / d64
        lo = max(0.0, np.max(np.minimum(t1, t2))); hi = min(1000.0, np.min(np.maximum(t1, t2)))
        if hi - lo > 1e-3:
            assert got is not None and abs(got[0] - lo) < 1e-4 * max(1, lo) and abs(got[1] - hi) < 1e-4 * max(1, hi)
        elif lo - hi > 1e-3:
            assert got is None


def test_moeller_trumbore_against_float64_plane_hit():
    """A textbook reduction: plane intersection + barycentric inside test in float64."""
    rng = np.random.default_rng(11)
    tris = rng.uniform(-5, 5, (400, 9)).astype(F32)
    sc = oracle.Scene(3, oracle.TRI_PLAIN, tris, build_tree=False)
    checked = 0
    for i in range(400):
        v0, v1, v2 = tris[i, :3].astype(np.float64), tris[i, 3:6].astype(np.float64), tris[i, 6:].astype(np.float64)
        ctr = (v0 + v1 + v2) / 3
        for _ in range(5):
            o = rng.uniform(-15, 15, 3).astype(F32)
            tgt = ctr + rng.uniform(-3, 3, 3)
            d = normalize_or_zero((tgt - o).astype(F32)[None])[0]
            n = np.cross(v1 - v0, v2 - v0)
            den = n @ d
            if abs(den) < 1e-3:
                continue
            t = n @ (v0 - o) / den
            p = o + t * d
            # barycentric in float64
            m = np.stack([v1 - v0, v2 - v0], 1)
            uv, *_ = np.linalg.lstsq(m, p - v0, rcond=None)
            inside = uv[0] >= 0 and uv[1] >= 0 and uv.sum() <= 1 and t > 1e-6
            margin = min(uv[0], uv[1], 1 - uv.sum())
            single = oracle.Scene(3, oracle.TRI_PLAIN, tris[i:i + 1])
            h = single.raycast_nearest(np.concatenate([o, d])[None], 1000.0, naive=True)
            if abs(margin) < 1e-4:
                continue
            checked += 1
            if inside:
                assert h["leaf"][0] == 0 and abs(h["t"][0] - t) < 1e-4 * max(1, abs(t))
            else:
                assert h["leaf"][0] == -1
    assert checked > 1000


def test_guard_only_drops_hits_outside_their_box():
    """Guarded and plain kinds agree except where the plain hit lies outside
    the primitive's own slab interval (so the guard cannot hide a real hit)."""
    from sq_amd import workloads as W
    sc = W.heightfield_scene(grid=257, extent=128.0)
    rays = W.incoherent_rays(sc, 100000)
    g = oracle.Scene(3, oracle.TRI, sc.prims)
    p = oracle.Scene(3, oracle.TRI_PLAIN, sc.prims)
    a = g.raycast_nearest(rays, 2000.0, threads=8)
    b = p.raycast_nearest(rays, 2000.0, threads=8)
    diff = (a["leaf"] != b["leaf"]) | (bits(a["t"]) != bits(b["t"]))
    assert diff.mean() < 1e-3
    # guarded BVH == guarded naive (the order-independence the device relies on)
    sub = rays[:3000]
    n = g.raycast_nearest(sub, 2000.0, naive=True)
    assert np.array_equal(n["leaf"], a["leaf"][:3000]) and np.array_equal(bits(n["t"]), bits(a["t"][:3000]))


# ---------------------------------------------------------------------------
# tree build and the required-work counter
# ---------------------------------------------------------------------------

def _total_key(f):
    i = np.asarray(f, F32).view(np.int32).astype(np.int64)
    return i ^ ((i >> 31) & 0x7FFFFFFF)


def _py_build(boxes):
    """tree.rs:112-163 from its prose, in Python: returns the node list."""
    dim = boxes.shape[1] // 2
    cent = ((boxes[:, :dim] + boxes[:, dim:]) * F32(0.5)).astype(F32)
    idx = list(range(len(boxes)))
    nodes = []

    def rec(start, end):
        sel = np.array(idx[start:end])
        b = np.concatenate([boxes[sel, :dim].min(0), boxes[sel, dim:].max(0)])
        me = len(nodes)
        nodes.append([b, start, 0, end - start])
        if end - start <= 4:
            return me
        cmin, cmax = cent[sel].min(0), cent[sel].max(0)
        ext = (cmax - cmin).astype(F32)
        axis = int(np.argmax(ext))          # first maximum
        if ext[axis] <= F32(1.1920929e-7):
            return me
        idx[start:end] = sorted(idx[start:end], key=lambda i: _total_key(cent[i, axis]))
        mid = start + (end - start) // 2
        a = rec(start, mid)
        c = rec(mid, end)
        nodes[me][1:] = [a, c, 0]
        return me

    if len(boxes):
        rec(0, len(boxes))
    return nodes, idx


@pytest.mark.parametrize("n,dim,seed", [(1, 3, 1), (5, 3, 2), (37, 2, 3), (200, 3, 4), (333, 6, 5)])
def test_build_matches_prose_restatement(n, dim, seed):
    s = Stream(seed)
    c = s.range(-10.0, 10.0, n * dim).reshape(n, dim)
    if n > 100:
        c[: n // 3] = c[0]                  # coincident centroids -> an over-full leaf
    h = s.range(0.1, 1.0, n * dim).reshape(n, dim)
    boxes = np.concatenate([c - h, c + h], 1).astype(F32)
    sc = oracle.Scene(dim, oracle.BOX, boxes)
    bounds, ca, cb, pc, idx = sc.export()
    nodes, pidx = _py_build(boxes)
    assert len(nodes) == len(ca)
    assert idx.tolist() == pidx
    for i, (b, a, c_, p) in enumerate(nodes):
        assert np.array_equal(bounds[i], b) and (ca[i], cb[i], pc[i]) == (a, c_, p)


def test_required_work_brute_force():
    """N_node = 1 + 2 (#interior nodes whose path passes), N_prim = prims of passing leaves."""
    s = Stream(77)
    boxes = np.concatenate([c := s.range(-20.0, 20.0, 300 * 3).reshape(300, 3), c + 1.0], 1).astype(F32)
    sc = oracle.Scene(3, oracle.BOX, boxes)
    bounds, ca, cb, pc, idx = sc.export()
    rays = np.concatenate([s.range(-30.0, 30.0, 200 * 3).reshape(200, 3),
                           normalize_or_zero(s.range(-1.0, 1.0, 200 * 3).reshape(200, 3))], 1).astype(F32)
    tl = s.range(1.0, 80.0, 200)
    nn, npr = sc.required_work(rays, tl)
    for r in range(200):
        want_n, want_p = 1, 0
        stack = [0]
        while stack:
            i = stack.pop()
            if oracle.ray_intersect(bounds[i], rays[r], float(tl[r])) is None:
                continue
            if pc[i]:
                want_p += int(pc[i])
            else:
                want_n += 2
                stack += [int(ca[i]), int(cb[i])]
        assert (nn[r], npr[r]) == (want_n, want_p)


# ---------------------------------------------------------------------------
# Box crossings (the AABB kind's test_ray_crossings): the builder's convention
# for a convex box, entry always (0 from inside), exit only when the box's own
# exit lies within max_toi.  Pinned by closed forms.
# ---------------------------------------------------------------------------

def _crossings(dim, box, o, d, max_toi):
    sc = oracle.Scene(dim, oracle.BOX, np.asarray([box], F32))
    ray = np.concatenate([np.asarray(o, F32), np.asarray(d, F32)])[None]
    off, leaf, t = sc.raycast_all(ray, max_toi, crossings=True)
    return leaf.tolist(), t.tolist()


def test_box_crossings_closed_forms():
    # box [1,3]^3 crossed along +x from x = 0: entry 1, exit 3
    assert _crossings(3, [1, 1, 1, 3, 3, 3], [0, 2, 2], [1, 0, 0], 100.0) == ([0, 0], [1.0, 3.0])
    # along -x from x = 5: entry 2, exit 4
    assert _crossings(3, [1, 1, 1, 3, 3, 3], [5, 2, 2], [-1, 0, 0], 100.0) == ([0, 0], [2.0, 4.0])
    # origin inside: entry 0 (the slab's t_min stays 0), exit 1
    assert _crossings(3, [1, 1, 1, 3, 3, 3], [2, 2, 2], [1, 0, 0], 100.0) == ([0, 0], [0.0, 1.0])
    # max_toi inside the box: the entry only (the exit lies beyond max_toi)
    assert _crossings(3, [1, 1, 1, 3, 3, 3], [0, 2, 2], [1, 0, 0], 2.0) == ([0], [1.0])
    # max_toi exactly at the exit: both
    assert _crossings(3, [1, 1, 1, 3, 3, 3], [0, 2, 2], [1, 0, 0], 3.0) == ([0, 0], [1.0, 3.0])
    # max_toi before the entry, or the ray beside the box: nothing
    assert _crossings(3, [1, 1, 1, 3, 3, 3], [0, 2, 2], [1, 0, 0], 0.5) == ([], [])
    assert _crossings(3, [1, 1, 1, 3, 3, 3], [0, 5, 2], [1, 0, 0], 100.0) == ([], [])
    # a diagonal ray through [0,4]^2 from (-1,-2): entry when y reaches 0 (t = 2 / |d|·...)
    d = np.array([1.0, 1.0], F32) / np.float32(np.sqrt(2.0))
    leaf, t = _crossings(2, [0, 0, 4, 4], [-1, -2], d, 100.0)
    inv = np.float32(1.0) / d[0]
    assert leaf == [0, 0]
    assert t[0] == float((np.float32(0) - np.float32(-2)) * inv)       # entry by y
    assert t[1] == float((np.float32(4) - np.float32(-1)) * inv)       # exit by x
    # D = 6: a box crossed along the last axis
    lo, hi = [0.0] * 6, [1.0] * 6
    lo[5], hi[5] = 10.0, 14.0
    assert _crossings(6, lo + hi, [0.5] * 5 + [0.0], [0, 0, 0, 0, 0, 1], 100.0) == ([0, 0], [10.0, 14.0])


def test_box_crossings_are_entry_and_exit_of_every_box():
    """On a random box scene every crossing list is the sorted union, over the
    boxes the walk reaches, of each box's slab entry and (when it lies within
    max_toi and after the entry) its exit, computed here independently from the
    slab closed form."""
    rng = np.random.default_rng(0xC2055)
    lo = rng.uniform(-20, 20, (300, 3)).astype(F32)
    boxes = np.concatenate([lo, lo + rng.uniform(0.5, 4.0, (300, 3)).astype(F32)], 1)
    sc = oracle.Scene(3, oracle.BOX, boxes)
    rays = np.concatenate([rng.uniform(-25, 25, (400, 3)), normalize_or_zero(rng.normal(size=(400, 3)).astype(F32))],
                          1).astype(F32)
    off, leaf, t = sc.raycast_all(rays, 30.0, crossings=True)
    for i in range(len(rays)):
        want = []
        for b in range(len(boxes)):
            hit = oracle.ray_intersect(boxes[b], rays[i], 30.0)
            if hit is None:
                continue
            te, _ = hit
            want.append((np.float32(te), b))
            full = oracle.ray_intersect(boxes[b], rays[i], np.inf)
            if full[1] > te and full[1] <= np.float32(30.0):
                want.append((np.float32(full[1]), b))
        want.sort()
        got = list(zip(t[off[i]:off[i + 1]].tolist(), leaf[off[i]:off[i + 1]].tolist()))
        assert got == [(float(a), b) for a, b in want], i


# ---------------------------------------------------------------------------
# Mesh instances (the builder's INST_MESH provider: test_ray_local is the
# mesh's own Bvh::raycast_nearest on the ray moved into the unscaled mesh
# frame, o / s and d / s, t kept).  Pinned by transforms under which every
# operation is exact, where the instance answer must equal the single-level
# tree over the explicitly transformed triangles and the naive scan.
# ---------------------------------------------------------------------------

def _dyadic_mesh(rng, n_tri):
    """Triangles with vertices on the 1/8 grid of [-4, 4]^3 (exact arithmetic
    for every product and sum of the walk and Moeller-Trumbore below)."""
    v = rng.integers(-32, 33, (n_tri, 9)).astype(F32) / F32(8.0)
    return v


def _signed_perm(rng):
    p = rng.permutation(3)
    s = rng.choice([-1.0, 1.0], 3)
    r = np.zeros((3, 3), F32)
    r[np.arange(3), p] = s
    return r


def test_mesh_instance_identity_equals_the_mesh_tree():
    """R = I, t = 0, s = 1: the instance answers exactly as the mesh's own
    single-level tree (nearest, triangle id, toi, normal), in both modes."""
    rng = np.random.default_rng(0x1D3)
    from sq_amd import workloads as W
    mesh = W.noisy_sphere(6, 0xA1)
    o = rng.uniform(-3, 3, (3000, 3)).astype(F32)
    tgt = rng.uniform(-0.9, 0.9, (3000, 3)).astype(F32)
    rays = np.concatenate([o, normalize_or_zero(tgt - o)], 1).astype(F32)
    for guarded in (False, True):
        inst = oracle.InstanceScene(3, oracle.INST_MESH, np.eye(3, dtype=F32)[None], np.zeros((1, 3), F32),
                                    mesh_tris=[mesh], mesh_of=np.zeros(1, np.uint32), scale=np.ones(1, F32),
                                    guarded=guarded)
        a = inst.raycast_nearest(rays, 100.0)
        b = oracle.Scene(3, oracle.TRI, mesh, guarded=guarded).raycast_nearest(rays, 100.0)
        hit = b["leaf"] >= 0
        assert hit.sum() > 1000
        assert np.array_equal(a["leaf"] >= 0, hit)
        assert np.array_equal(a["sub"][hit], b["leaf"][hit])
        assert np.array_equal(bits(a["t"]), bits(b["t"]))
        assert np.array_equal(bits(a["normal"][hit]), bits(b["normal"][hit]))


def test_mesh_instances_dyadic_transforms_equal_world_triangles():
    """Signed-permutation rotations, power-of-two scales, translations and
    vertices on a dyadic grid, and dyadic unnormalised ray directions: every
    operation of the instance test (R^T(o - t), d / s, Moeller-Trumbore in the
    mesh frame) is exact, so the two-level answer equals, bit for bit, the
    single-level tree over the triangles transformed explicitly into the world
    (R v s + t) and the naive scan over them (guarded mode, whose answer does
    not depend on the tree).  A transposed rotation, a scale applied the wrong
    way or a dropped translation fails it."""
    rng = np.random.default_rng(0xD1AD)
    meshes = [_dyadic_mesh(rng, 60), _dyadic_mesh(rng, 90)]
    n_inst = 40
    rot = np.stack([_signed_perm(rng) for _ in range(n_inst)])
    scale = rng.choice([0.5, 1.0, 2.0, 4.0], n_inst).astype(F32)
    trans = (rng.integers(-160, 161, (n_inst, 3)) / 4.0).astype(F32)
    mesh_of = rng.integers(0, 2, n_inst).astype(np.uint32)
    world, owner = [], []
    for i in range(n_inst):
        m = meshes[mesh_of[i]].reshape(-1, 3, 3)
        w = np.einsum("ab,tvb->tva", rot[i], m) * scale[i] + trans[i]
        world.append(w.reshape(-1, 9).astype(F32))
        owner += [(i, k) for k in range(len(m))]
    world = np.concatenate(world)
    o = (rng.integers(-200, 201, (6000, 3)) / 4.0).astype(F32)
    d = (rng.integers(-4, 5, (6000, 3)) / 4.0).astype(F32)
    keep = np.any(d != 0, axis=1)
    rays = np.concatenate([o[keep], d[keep]], 1)
    inst = oracle.InstanceScene(3, oracle.INST_MESH, rot, trans, mesh_tris=meshes, mesh_of=mesh_of, scale=scale,
                                guarded=True)
    a = inst.raycast_nearest(rays, 1000.0)
    flat = oracle.Scene(3, oracle.TRI, world, guarded=True)
    b = flat.raycast_nearest(rays, 1000.0)
    nv = flat.raycast_nearest(rays, 1000.0, naive=True)
    hit = b["leaf"] >= 0
    assert hit.sum() > 300
    assert np.array_equal(b["leaf"], nv["leaf"]) and np.array_equal(bits(b["t"]), bits(nv["t"]))
    assert np.array_equal(a["leaf"] >= 0, hit)
    got = [(int(a["leaf"][j]), int(a["sub"][j])) for j in np.nonzero(hit)[0]]
    want = [owner[b["leaf"][j]] for j in np.nonzero(hit)[0]]
    # the same triangle answers; ties between instances sharing a toi are
    # ordered by instance id on the two-level side and by global index on the
    # flat side, which rank instances the same way (owner is instance-major)
    assert got == want
    assert np.array_equal(bits(a["t"]), bits(b["t"]))
    # all hits (the same set in both modes): one per instance, its nearest
    # triangle, equal to the per-instance minimum of the flat all-hits list
    off, leaf, t, sub = inst.raycast_all(rays[:500], 1000.0)
    foff, fleaf, ft = flat.raycast_all(rays[:500], 1000.0)
    for j in range(500):
        best = {}
        for g, tt in zip(fleaf[foff[j]:foff[j + 1]], ft[foff[j]:foff[j + 1]]):
            i, k = owner[g]
            if i not in best or (tt, k) < best[i]:
                best[i] = (tt, k)
        want = sorted((tt, i, k) for i, (tt, k) in best.items())
        got = sorted(zip(t[off[j]:off[j + 1]].tolist(), leaf[off[j]:off[j + 1]].tolist(), sub[off[j]:off[j + 1]].tolist()))
        assert [(float(a_), int(b_), int(c_)) for a_, b_, c_ in want] == got, j
