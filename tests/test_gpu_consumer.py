"""Consumer narrow phases compiled at run time (sq_narrow_create, hipRTC) into
the nearest-hit walk, against the CPU oracle running the same source.

  * the built-in ball as a consumer (D = 3 and 5) equals the engine's built-in
    SPHERE kind bit for bit, and the oracle;
  * a consumer ellipsoid equals the oracle's consumer kind (reference LIFO
    walk, same source built by g++) bit for bit;
  * a source that does not compile reports the compiler's log; a consumer tree
    refuses the built-in queries; an empty tree answers misses;
  * all hits and crossings (pooled collect through the consumer): the ball
    consumer's all-hits equal the built-in SPHERE's bit for bit (and its
    crossings, with no crossings function, equal its all-hits); the shells of
    crossings.rs and the marched implicit blobs of field.rs give the oracle's
    consumer crossings (same source built by g++) bit for bit, CSR offsets,
    leaf order and toi, and the blobs' nearest hits equal the oracle's;
  * shape casts and overlap through the consumer (sq_narrow_cast /
    sq_narrow_overlap): the ball consumer equals the built-in SPHERE kind and
    the oracle bit for bit; a source without them finds no contact and
    overlaps at the box level, as the oracle.
"""
import numpy as np
import pytest

import oracle
import sq_amd
from sq_amd import workloads as W

from consumer_narrow import (BALL, BALL_FULL, BLOB, ELLIPSOID, SHELL, cast_overlap_entry_points, compile_cpu,
                             cpu_entry_points)
from helpers import assert_nearest_equal

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def host(t):
    return t.cpu().numpy()


def balls(n, dim, seed, half):
    s = W.Stream(seed)
    c = s.range(-half, half, n * dim).reshape(n, dim)
    r = s.range(0.3, 3.0, n).reshape(n, 1)
    return np.concatenate([c, r], 1).astype(np.float32)


@pytest.mark.parametrize("dim,half", [(3, 40.0), (5, 10.0)])
def test_consumer_ball_equals_builtin_sphere(ctx, tmp_path, dim, half):
    print(f"case: consumer ball D={dim}", flush=True)
    b = balls(20000, dim, 0xD0 + dim, half)
    boxes = np.concatenate([b[:, :dim] - b[:, dim:], b[:, :dim] + b[:, dim:]], 1).astype(np.float32)
    rays = W.random_rays(20000, 0xD1 + dim, -half - 5.0, half + 5.0, dim=dim)
    nw = sq_amd.Narrow(ctx, dim, BALL, dim + 1)
    cb = nw.build(boxes, b)
    got = cb.raycast_nearest(dev(rays), 300.0)
    builtin = sq_amd.DeviceBvh.build(ctx, dim, sq_amd.SPHERE, b).raycast_nearest(dev(rays), 300.0)
    assert np.array_equal(host(got["leaf"]), host(builtin["leaf"]))
    assert np.array_equal(host(got["toi"]).view(np.uint32), host(builtin["toi"]).view(np.uint32))
    lib, addr = compile_cpu(BALL, dim, dim + 1, str(tmp_path))
    ref = oracle.Scene.consumer(dim, boxes, b, addr).raycast_nearest(rays, 300.0, threads=8)
    assert (ref["leaf"] >= 0).sum() > 1000
    assert_nearest_equal(host(got["leaf"]), host(got["toi"]), ref["leaf"], ref["t"], f"consumer ball D={dim}")
    cb.close()
    nw.close()


def test_consumer_hits_outside_their_boxes(ctx, tmp_path):
    """A provider whose hits leave their own boxes: balls whose world boxes are
    shrunk to 0.6 of the radius (the reference asks only that the ray meet the
    leaf box, geometry.rs:37-45).  Here the walk order decides many answers;
    the default mode returns the reference's LIFO walk's answer bit for bit
    (oracle consumer kind, same source built by g++), which differs from the
    naive scan and from the guarded answer on many rays; SQ_HITS_GUARDED
    returns the guarded oracle's."""
    print("case: consumer hits outside their boxes, both modes", flush=True)
    dim = 3
    b = balls(20000, dim, 0xD8, 40.0)
    shrink = b[:, dim:] * np.float32(0.6)
    boxes = np.concatenate([b[:, :dim] - shrink, b[:, :dim] + shrink], 1).astype(np.float32)
    rays = W.random_rays(30000, 0xD9, -45.0, 45.0, dim=dim)
    nw = sq_amd.Narrow(ctx, dim, BALL, dim + 1)
    cb = nw.build(boxes, b)
    lib, addr = compile_cpu(BALL, dim, dim + 1, str(tmp_path))
    sc = oracle.Scene.consumer(dim, boxes, b, addr)
    ref = sc.raycast_nearest(rays, 300.0, threads=8)
    naive = sc.raycast_nearest(rays, 300.0, naive=True)
    guarded = oracle.Scene.consumer(dim, boxes, b, addr, guarded=True).raycast_nearest(rays, 300.0, threads=8)
    got = cb.raycast_nearest(dev(rays), 300.0)
    assert_nearest_equal(host(got["leaf"]), host(got["toi"]), ref["leaf"], ref["t"], "consumer, reference order")
    order = int((ref["leaf"] != naive["leaf"]).sum())
    gdiff = int((ref["leaf"] != guarded["leaf"]).sum())
    print(f"rays where order decides (reference walk != naive scan): {order}; guarded differs on {gdiff}", flush=True)
    assert order > 10 and gdiff > 10
    try:
        ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED)
        g = cb.raycast_nearest(dev(rays), 300.0)
        assert_nearest_equal(host(g["leaf"]), host(g["toi"]), guarded["leaf"], guarded["t"], "consumer, guarded mode")
    finally:
        ctx.set_hit_mode(sq_amd.SQ_HITS_REFERENCE)
    cb.close()
    nw.close()


def test_consumer_ellipsoid_matches_oracle(ctx, tmp_path):
    print("case: consumer ellipsoid", flush=True)
    s = W.Stream(0xD5)
    c = s.range(-60.0, 60.0, 30000 * 3).reshape(-1, 3)
    a = s.range(0.2, 3.0, 30000 * 3).reshape(-1, 3)
    prims = np.concatenate([c, a], 1).astype(np.float32)
    boxes = np.concatenate([c - a, c + a], 1).astype(np.float32)
    rays = np.concatenate([W.random_rays(30000, 0xD6, -70.0, 70.0),
                           W.random_rays(2000, 0xD7, -5.0, 5.0)])
    nw = sq_amd.Narrow(ctx, 3, ELLIPSOID, 6)
    cb = nw.build(boxes, prims)
    got = cb.raycast_nearest(dev(rays), 500.0)
    lib, addr = compile_cpu(ELLIPSOID, 3, 6, str(tmp_path))
    ref = oracle.Scene.consumer(3, boxes, prims, addr).raycast_nearest(rays, 500.0, threads=8)
    assert (ref["leaf"] >= 0).sum() > 3000
    assert_nearest_equal(host(got["leaf"]), host(got["toi"]), ref["leaf"], ref["t"], "consumer ellipsoid")
    # the tree is the reference median split over the consumer's boxes
    st = cb.stats()
    ts = sq_amd.Tree.build(3, sq_amd.AABB, boxes).stats()
    for k in ("node_count", "leaf_count", "max_depth"):
        assert st[k] == ts[k]
    cb.close()
    nw.close()


def test_consumer_errors_and_empty_tree(ctx):
    print("case: consumer compile error, refusals, empty tree", flush=True)
    with pytest.raises(sq_amd.SqError) as e:
        sq_amd.Narrow(ctx, 3, "__device__ bool sq_narrow(const float* p) { return undefined_name; }", 4)
    assert "error" in str(e.value)
    nw = sq_amd.Narrow(ctx, 3, BALL, 4)
    empty = nw.build(np.zeros((0, 6), np.float32), np.zeros((0, 4), np.float32))
    r = empty.raycast_nearest(dev(W.random_rays(100, 1, -5.0, 5.0)), 100.0)
    assert (host(r["leaf"]) == -1).all() and np.isinf(host(r["toi"])).all()
    # a consumer tree answers only through its narrow phase
    b = balls(50, 3, 0xD8, 10.0)
    boxes = np.concatenate([b[:, :3] - b[:, 3:], b[:, :3] + b[:, 3:]], 1).astype(np.float32)
    cb = nw.build(boxes, b)
    bogus = sq_amd.DeviceBvh.__new__(sq_amd.DeviceBvh)
    bogus._h, bogus.ctx, bogus.dim, bogus.kind, bogus.n_prims = cb._h, ctx, 3, sq_amd.AABB, 50
    with pytest.raises(sq_amd.SqError):
        bogus.raycast_nearest(dev(W.random_rays(10, 2, -5.0, 5.0)), 100.0)
    bogus._h = None
    other = sq_amd.Narrow(ctx, 3, ELLIPSOID, 6)
    wrong = sq_amd.ConsumerBvh(other, cb._h, 50)        # the tree under another narrow phase
    with pytest.raises(sq_amd.SqError):
        wrong.raycast_nearest(dev(W.random_rays(10, 3, -5.0, 5.0)), 100.0)
    wrong._h = None
    cb.close()
    empty.close()
    other.close()
    nw.close()


def csr_host(off, leaf, toi):
    return host(off).astype(np.uint64), host(leaf).astype(np.int64), host(toi)


def assert_csr_equal(got, ref, what):
    go, gl, gt = got
    ro, rl, rt = ref
    assert np.array_equal(go, ro.astype(np.uint64)), f"{what}: offsets differ"
    assert np.array_equal(gl, rl), f"{what}: leaf ids differ"
    assert np.array_equal(gt.view(np.uint32), rt.view(np.uint32)), f"{what}: toi differs"


def test_consumer_all_hits_ball_equals_builtin(ctx):
    print("case: consumer ball all-hits vs built-in SPHERE", flush=True)
    b = balls(20000, 3, 0xDA, 40.0)
    boxes = np.concatenate([b[:, :3] - b[:, 3:], b[:, :3] + b[:, 3:]], 1).astype(np.float32)
    rays = dev(W.random_rays(20000, 0xDB, -45.0, 45.0))
    nw = sq_amd.Narrow(ctx, 3, BALL, 4)
    cb = nw.build(boxes, b)
    got = csr_host(*cb.raycast_all(rays, 300.0))
    builtin = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.SPHERE, b)
    ref = csr_host(*builtin.raycast_all(rays, 300.0))
    assert ref[1].shape[0] > 5000
    assert_csr_equal(got, ref, "ball all-hits")
    # no crossings function: the crossings are the test_ray hits
    assert_csr_equal(csr_host(*cb.raycast_all(rays, 300.0, crossings=True)), ref, "ball crossings")
    # a capacity too small is reported and the Python side retries with the total
    small = csr_host(*cb.raycast_all(rays, 300.0, capacity=16))
    assert_csr_equal(small, ref, "ball all-hits from capacity 16")
    cb.close()
    nw.close()


@pytest.mark.parametrize("dim", [3, 4])
def test_consumer_shell_crossings_match_oracle(ctx, tmp_path, dim):
    print(f"case: consumer shell crossings D={dim}", flush=True)
    s = W.Stream(0xDC + dim)
    n = 30000
    c = s.range(-60.0, 60.0, n * dim).reshape(-1, dim)
    r = s.range(1.0, 4.0, n).reshape(-1, 1)
    sh = np.concatenate([c, r], 1).astype(np.float32)
    boxes = np.concatenate([c - r, c + r], 1).astype(np.float32)
    rays = W.random_rays(8000, 0xDD + dim, -65.0, 65.0, dim=dim)
    nw = sq_amd.Narrow(ctx, dim, SHELL, dim + 1)
    cb = nw.build(boxes, sh)
    got = csr_host(*cb.raycast_all(dev(rays), 500.0, crossings=True))
    lib, addr, xaddr = cpu_entry_points(SHELL, dim, dim + 1, str(tmp_path))
    sc = oracle.Scene.consumer(dim, boxes, sh, addr, xaddr)
    ref = sc.raycast_all(rays, 500.0, crossings=True)
    assert ref[1].shape[0] > 3000
    assert_csr_equal(got, ref, f"shell crossings D={dim}")
    # nearest: the first root at or after the origin
    near = cb.raycast_nearest(dev(rays), 500.0)
    rn = sc.raycast_nearest(rays, 500.0, threads=8)
    assert_nearest_equal(host(near["leaf"]), host(near["toi"]), rn["leaf"], rn["t"], f"shell nearest D={dim}")
    cb.close()
    nw.close()


def test_consumer_blob_crossings_match_oracle(ctx, tmp_path):
    print("case: consumer implicit blobs (marched field)", flush=True)
    s = W.Stream(0xDE)
    n = 400
    c = s.range(-40.0, 40.0, n * 3).reshape(n, 3)
    r0 = s.range(2.0, 5.0, n).reshape(n, 1)
    amp = s.range(0.0, 2.0, n).reshape(n, 1)
    freq = s.range(0.5, 1.5, n).reshape(n, 1)
    prims = np.concatenate([c, r0, amp, freq], 1).astype(np.float32)
    h = (r0 + amp + 1.0).astype(np.float32)
    boxes = np.concatenate([c - h, c + h], 1).astype(np.float32)
    rays = W.random_rays(3000, 0xDF, -45.0, 45.0)
    nw = sq_amd.Narrow(ctx, 3, BLOB, 6)
    cb = nw.build(boxes, prims)
    lib, addr, xaddr = cpu_entry_points(BLOB, 3, 6, str(tmp_path))
    sc = oracle.Scene.consumer(3, boxes, prims, addr, xaddr)
    got = csr_host(*cb.raycast_all(dev(rays), 120.0, crossings=True))
    ref = sc.raycast_all(rays, 120.0, crossings=True)
    assert ref[1].shape[0] > 1000
    assert_csr_equal(got, ref, "blob crossings")
    near = cb.raycast_nearest(dev(rays), 120.0)
    rn = sc.raycast_nearest(rays, 120.0, threads=8)
    assert (rn["leaf"] >= 0).sum() > 300
    assert_nearest_equal(host(near["leaf"]), host(near["toi"]), rn["leaf"], rn["t"], "blob nearest")
    cb.close()
    nw.close()


def test_consumer_cast_and_overlap(ctx, tmp_path):
    print("case: consumer shape casts and overlap", flush=True)
    b = balls(30000, 3, 0xE0, 60.0)
    boxes = np.concatenate([b[:, :3] - b[:, 3:], b[:, :3] + b[:, 3:]], 1).astype(np.float32)
    rays = W.random_rays(20000, 0xE1, -65.0, 65.0)
    casts = np.concatenate([rays, np.full((20000, 3), -0.75, np.float32), np.full((20000, 3), 0.75, np.float32)], 1)
    s = W.Stream(0xE2)
    c = s.range(-60.0, 60.0, 20000 * 3).reshape(-1, 3)
    h = s.range(0.5, 6.0, 20000 * 3).reshape(-1, 3)
    regions = np.concatenate([c - h, c + h], 1).astype(np.float32)
    nw = sq_amd.Narrow(ctx, 3, BALL_FULL, 4)
    cb = nw.build(boxes, b)
    builtin = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.SPHERE, b)
    got = cb.shapecast_nearest(dev(casts), 300.0)
    want = builtin.shapecast_nearest(dev(casts), 300.0)
    assert (host(want["leaf"]) >= 0).sum() > 2000
    assert np.array_equal(host(got["leaf"]), host(want["leaf"]))
    assert np.array_equal(host(got["toi"]).view(np.uint32), host(want["toi"]).view(np.uint32))
    ga = csr_host(*cb.shapecast_all(dev(casts), 300.0))
    wa = csr_host(*builtin.shapecast_all(dev(casts), 300.0))
    assert wa[1].shape[0] > 3000
    assert_csr_equal(ga, wa, "consumer cast all vs SPHERE")
    go = cb.overlap(dev(regions))
    wo = builtin.overlap(dev(regions))
    assert np.array_equal(host(go[0]), host(wo[0])) and np.array_equal(host(go[1]), host(wo[1]))
    assert host(wo[1]).shape[0] > 3000
    # the oracle's consumer kind running the same source
    lib, addr, xaddr = cpu_entry_points(BALL_FULL, 3, 4, str(tmp_path))
    caddr, oaddr = cast_overlap_entry_points(lib)
    sc = oracle.Scene.consumer(3, boxes, b, addr, xaddr, caddr, oaddr)
    rn = sc.shapecast_nearest(casts, 300.0)
    assert_nearest_equal(host(got["leaf"]), host(got["toi"]), rn["leaf"], rn["t"], "consumer cast vs oracle")
    ro = sc.overlap(regions)
    assert np.array_equal(host(go[0]).astype(np.uint64), ro[0].astype(np.uint64))
    assert np.array_equal(host(go[1]).astype(np.int64), ro[1])
    cb.close()
    nw.close()
    # without sq_narrow_cast / sq_narrow_overlap: no contact, box-level overlap
    nw2 = sq_amd.Narrow(ctx, 3, BALL, 4)
    cb2 = nw2.build(boxes, b)
    miss = cb2.shapecast_nearest(dev(casts), 300.0)
    assert (host(miss["leaf"]) == -1).all()
    assert int(host(cb2.shapecast_all(dev(casts), 300.0)[0])[-1]) == 0
    bo = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.AABB, boxes).overlap(dev(regions))
    po = cb2.overlap(dev(regions))
    assert np.array_equal(host(po[0]), host(bo[0])) and np.array_equal(host(po[1]), host(bo[1]))
    cb2.close()
    nw2.close()
