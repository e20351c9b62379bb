"""The oracle's consumer kind (a narrow phase through a C function, compiled
from the same source the engine compiles with hipRTC), pinned on the CPU:

  * the built-in ball written as a consumer gives the built-in SPHERE kind's
    results bit for bit, in D = 3 and D = 5 (same arithmetic, same guard);
  * with a consumer ellipsoid, the oracle's BVH walk equals its naive scan
    (the reference's BVH == naive contract, sq/src/lib.rs:15-17).
"""
import numpy as np

import oracle
from sq_amd import workloads as W

from consumer_narrow import BALL, ELLIPSOID, compile_cpu


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
