"""The oracle's consumer kind (a narrow phase through a C function, compiled
from the same source the engine compiles with hipRTC), pinned on the CPU:

  * the built-in ball written as a consumer gives the built-in SPHERE kind's
    results bit for bit, in D = 3 and D = 5 (same arithmetic, same guard);
  * with a consumer ellipsoid, the oracle's BVH walk equals its naive scan
    (the reference's BVH == naive contract, sq/src/lib.rs:15-17);
  * consumer crossings (test_ray_crossings through a C function): the shells
    of crossings.rs written as a consumer cross a radius-2 shell at 8 and 12
    (crossings.rs:113-135) and give the built-in SHELL kind's crossings except
    the few the box guard drops; the implicit blobs of field.rs converge to
    the analytic sphere at amp 0 (field.rs:178-214) and their BVH crossings
    and nearest hits equal the naive scan (field.rs:160-176);
  * consumer shape casts and overlap (test_shape_cast / test_overlap through
    C functions): the ball written as a consumer casts and overlaps like the
    built-in SPHERE kind bit for bit; without them a cast finds nothing and
    overlap is the box test.
"""
import numpy as np

import oracle
from sq_amd import workloads as W

from consumer_narrow import (BALL, BALL_FULL, BLOB, ELLIPSOID, SHELL, cast_overlap_entry_points, compile_cpu,
                             cpu_entry_points)


def balls(n, dim, seed, half=40.0):
    s = W.Stream(seed)
    c = s.range(-half, half, n * dim).reshape(n, dim)
    r = s.range(0.3, 3.0, n).reshape(n, 1)
    return np.concatenate([c, r], 1).astype(np.float32)


def ball_boxes(b, dim):
    c, r = b[:, :dim], b[:, dim:dim + 1]
    return np.concatenate([c - r, c + r], 1).astype(np.float32)


def test_consumer_ball_equals_builtin_sphere(tmp_path):
    for dim, half in ((3, 40.0), (5, 10.0)):
        lib, addr = compile_cpu(BALL, dim, dim + 1, str(tmp_path))
        b = balls(3000, dim, 0xC0 + dim, half)
        rays = W.random_rays(3000, 0xC1 + dim, -half - 5.0, half + 5.0, dim=dim)
        ref = oracle.Scene(dim, oracle.SPHERE, b).raycast_nearest(rays, 300.0)
        got = oracle.Scene.consumer(dim, ball_boxes(b, dim), b, addr).raycast_nearest(rays, 300.0)
        assert (ref["leaf"] >= 0).sum() > 100
        assert np.array_equal(got["leaf"], ref["leaf"]), f"D={dim}: ids differ"
        assert np.array_equal(got["t"].view(np.uint32), ref["t"].view(np.uint32)), f"D={dim}: toi differs"
        del lib


def test_consumer_ellipsoid_bvh_equals_naive(tmp_path):
    lib, addr = compile_cpu(ELLIPSOID, 3, 6, str(tmp_path))
    s = W.Stream(0xE1)
    c = s.range(-40.0, 40.0, 2000 * 3).reshape(-1, 3)
    a = s.range(0.3, 4.0, 2000 * 3).reshape(-1, 3)
    prims = np.concatenate([c, a], 1).astype(np.float32)
    boxes = np.concatenate([c - a, c + a], 1).astype(np.float32)
    rays = W.random_rays(4000, 0xE2, -45.0, 45.0)
    sc = oracle.Scene.consumer(3, boxes, prims, addr)
    bvh = sc.raycast_nearest(rays, 300.0)
    naive = sc.raycast_nearest(rays, 300.0, naive=True)
    assert (bvh["leaf"] >= 0).sum() > 200
    assert np.array_equal(bvh["leaf"], naive["leaf"])
    assert np.array_equal(bvh["t"].view(np.uint32), naive["t"].view(np.uint32))
    del lib


def ray_new(o, d):
    """Ray::new (ray.rs:15-36): the direction normalised as point.rs:60-89."""
    return np.concatenate([np.asarray(o, np.float32), W.normalize_or_zero(np.asarray(d, np.float32)[None])[0]])[None, :]


def lists(off, leaf, t):
    return [(leaf[off[i]:off[i + 1]].tolist(), t[off[i]:off[i + 1]].view(np.uint32).tolist()) for i in range(len(off) - 1)]


def test_consumer_shell_crossings_entry_exit(tmp_path):
    lib, addr, xaddr = cpu_entry_points(SHELL, 3, 4, str(tmp_path))
    assert xaddr
    sh = np.array([[10, 0, 0, 2]], np.float32)
    sc = oracle.Scene.consumer(3, ball_boxes(sh, 3), sh, addr, xaddr)
    off, leaf, t = sc.raycast_all(ray_new([0, 0, 0], [1, 0, 0]), 100.0, crossings=True)
    assert off.tolist() == [0, 2] and leaf.tolist() == [0, 0] and t.tolist() == [8.0, 12.0]
    off, leaf, t = sc.raycast_all(ray_new([0, 0, 0], [1, 0, 0]), 100.0)
    assert leaf.tolist() == [0] and t.tolist() == [8.0]
    del lib


def test_consumer_shell_crossings_equal_builtin_shell(tmp_path):
    for dim in (3, 4):
        lib, addr, xaddr = cpu_entry_points(SHELL, dim, dim + 1, str(tmp_path))
        s = W.Stream(0x5E11 + dim)
        c = s.range(-30.0, 30.0, 1500 * dim).reshape(-1, dim)
        r = s.range(1.0, 4.0, 1500).reshape(-1, 1)
        sh = np.concatenate([c, r], 1).astype(np.float32)
        rays = W.random_rays(3000, 0x5E21 + dim, -35.0, 35.0, dim=dim)
        ref = lists(*oracle.Scene(dim, oracle.SHELL, sh).raycast_all(rays, 500.0, crossings=True))
        cons = oracle.Scene.consumer(dim, ball_boxes(sh, dim), sh, addr, xaddr)
        got = lists(*cons.raycast_all(rays, 500.0, crossings=True))
        assert sum(len(x[0]) for x in ref) > 800
        differ = 0
        for (gl, gt), (rl, rt) in zip(got, ref):
            if (gl, gt) == (rl, rt):
                continue
            differ += 1
            # the guard only drops crossings (outside the box's slab interval)
            it = iter(zip(rl, rt))
            assert all(any(e == x for x in it) for e in zip(gl, gt))
        assert differ <= len(ref) // 500, f"D={dim}: {differ} rays differ"
        # the consumer's BVH walk equals its naive scan
        nv = lists(*cons.raycast_all(rays, 500.0, crossings=True, naive=True))
        assert got == nv
        del lib


def test_consumer_blob_converges_to_the_sphere(tmp_path):
    lib, addr, xaddr = cpu_entry_points(BLOB, 3, 6, str(tmp_path))
    blob = np.array([[20, 0, 0, 3, 0, 1]], np.float32)
    h = 3 + 0 + 1
    box = np.array([[20 - h, -h, -h, 20 + h, h, h]], np.float32)
    sc = oracle.Scene.consumer(3, box, blob, addr, xaddr)
    ray = ray_new([0, 0, 0], [1, 0, 0])
    near = sc.raycast_nearest(ray, 100.0)
    assert near["leaf"][0] == 0 and abs(near["t"][0] - 17.0) < 1e-2
    off, leaf, t = sc.raycast_all(ray, 100.0, crossings=True)
    assert off.tolist() == [0, 2]
    assert abs(t[0] - 17.0) < 1e-2 and abs(t[1] - 23.0) < 1e-2
    # a ray that misses every box, and the sine polynomial itself
    assert sc.raycast_nearest(ray_new([0, 50, 0], [1, 0, 0]), 100.0)["leaf"][0] == -1
    del lib


def blobs(n, seed):
    s = W.Stream(seed)
    c = s.range(-40.0, 40.0, n * 3).reshape(n, 3)
    r0 = s.range(2.0, 5.0, n).reshape(n, 1)
    amp = s.range(0.0, 2.0, n).reshape(n, 1)
    freq = s.range(0.5, 1.5, n).reshape(n, 1)
    prims = np.concatenate([c, r0, amp, freq], 1).astype(np.float32)
    h = (r0 + amp + 1.0).astype(np.float32)
    return np.concatenate([c - h, c + h], 1).astype(np.float32), prims


def aimed_rays(prims, n, seed):
    """Origins U[-60, 60]^3, each aimed at a blob centre (offset up to 3
    units), so most rays cross a blob."""
    s = W.Stream(seed)
    o = s.range(-60.0, 60.0, n * 3).reshape(n, 3)
    which = (s.u64(n) % np.uint64(prims.shape[0])).astype(np.int64)
    target = prims[which, :3] + s.range(-3.0, 3.0, n * 3).reshape(n, 3)
    d = W.normalize_or_zero((target - o).astype(np.float32))
    return np.concatenate([o, d], 1).astype(np.float32)


def test_consumer_blob_bvh_equals_naive(tmp_path):
    lib, addr, xaddr = cpu_entry_points(BLOB, 3, 6, str(tmp_path))
    for k in range(4):
        boxes, prims = blobs(40, 0xF1E1 + k)
        rays = aimed_rays(prims, 60, 0xF1F1 + k)
        sc = oracle.Scene.consumer(3, boxes, prims, addr, xaddr)
        a, b = sc.raycast_nearest(rays, 150.0), sc.raycast_nearest(rays, 150.0, naive=True)
        assert np.array_equal(a["leaf"], b["leaf"]) and np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
        ca = lists(*sc.raycast_all(rays, 150.0, crossings=True))
        cb = lists(*sc.raycast_all(rays, 150.0, crossings=True, naive=True))
        assert ca == cb
        assert sum(len(x[0]) for x in ca) > 40
    del lib


def ball_casts(n, seed, half, probe=0.75):
    rays = W.random_rays(n, seed, -half - 5.0, half + 5.0)
    return np.concatenate([rays, np.full((n, 3), -probe, np.float32), np.full((n, 3), probe, np.float32)], 1)


def test_consumer_cast_and_overlap_equal_builtin_sphere(tmp_path):
    lib, addr, xaddr = cpu_entry_points(BALL_FULL, 3, 4, str(tmp_path))
    caddr, oaddr = cast_overlap_entry_points(lib)
    assert caddr and oaddr
    b = balls(3000, 3, 0xC7, 40.0)
    cons = oracle.Scene.consumer(3, ball_boxes(b, 3), b, addr, xaddr, caddr, oaddr)
    ref = oracle.Scene(3, oracle.SPHERE, b)
    casts = ball_casts(3000, 0xC8, 40.0)
    a, r = cons.shapecast_nearest(casts, 300.0), ref.shapecast_nearest(casts, 300.0)
    assert (r["leaf"] >= 0).sum() > 300
    assert np.array_equal(a["leaf"], r["leaf"]) and np.array_equal(a["t"].view(np.uint32), r["t"].view(np.uint32))
    ao, al, at = cons.shapecast_all(casts[:500], 300.0)
    ro, rl, rt = ref.shapecast_all(casts[:500], 300.0)
    assert np.array_equal(ao, ro) and np.array_equal(al, rl) and np.array_equal(at.view(np.uint32), rt.view(np.uint32))
    s = W.Stream(0xC9)
    c = s.range(-40.0, 40.0, 2000 * 3).reshape(-1, 3)
    h = s.range(0.5, 6.0, 2000 * 3).reshape(-1, 3)
    reg = np.concatenate([c - h, c + h], 1).astype(np.float32)
    co, cl = cons.overlap(reg)
    so, sl = ref.overlap(reg)
    assert len(sl) > 500 and np.array_equal(co, so) and np.array_equal(cl, sl)
    # a source without them: no contact, and overlap is the box test
    lib2, addr2 = compile_cpu(BALL, 3, 4, str(tmp_path))
    plain = oracle.Scene.consumer(3, ball_boxes(b, 3), b, addr2)
    assert (plain.shapecast_nearest(casts, 300.0)["leaf"] == -1).all()
    boxes = oracle.Scene(3, oracle.BOX, ball_boxes(b, 3))
    po, pl = plain.overlap(reg)
    bo, bl = boxes.overlap(reg)
    assert np.array_equal(po, bo) and np.array_equal(pl, bl) and len(bl) > len(sl)
    del lib, lib2
