"""Ball instances (SQ_INST_BALL: the reference's BallInstances, shape_overlap.rs:
288-345) and the two-level shape casts (Tlas::shapecast_nearest / _all,
tlas.rs:306-391), against the CPU oracle's Tlas over the same instances:

  * rays: nearest (instance, toi and world normal bit for bit), any-hit, all
    hits, crossings (one per instance, the trait default);
  * overlap with the balls' exact test (distance from the centre <= r);
  * shape casts with a ball probe, nearest and all, CSR bit for bit;
  * box instances answer no cast (the trait default); a batch set over the
    ball TLAS equals its single calls.
"""
import numpy as np
import pytest

import oracle
import sq_amd
from sq_amd import workloads as W

from helpers import assert_nearest_equal

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def host(t):
    return t.cpu().numpy()


def ball_scene(n, seed, rotate):
    s = W.Stream(seed)
    rot = W.random_rotations(s, n) if rotate else np.tile(np.eye(3, dtype=np.float32), (n, 1, 1))
    trans = s.range(-200.0, 200.0, n * 3).reshape(n, 3).astype(np.float32)
    radii = s.range(0.5, 4.0, n).astype(np.float32)
    return np.ascontiguousarray(rot, np.float32), trans, radii


def casts_of(rays, probe):
    n = rays.shape[0]
    return np.concatenate([rays[:, :6], np.full((n, 3), -probe, np.float32), np.full((n, 3), probe, np.float32)], 1)


def assert_csr(got, ref, what, has_t=True):
    off, leaf = host(got[0]).astype(np.uint64), host(got[1]).astype(np.int64)
    assert np.array_equal(off, ref[0].astype(np.uint64)), f"{what}: offsets"
    assert np.array_equal(leaf, ref[1]), f"{what}: ids"
    if has_t:
        assert np.array_equal(host(got[2]).view(np.uint32), ref[2].view(np.uint32)), f"{what}: toi"


@pytest.mark.parametrize("rotate", [False, True])
def test_ball_instances_match_oracle(ctx, rotate):
    print(f"case: ball instances, rotated={rotate}", flush=True)
    rot, trans, radii = ball_scene(20000, 0xBA10 + rotate, rotate)
    tl = sq_amd.Tlas(ctx, 3, sq_amd.Tlas.BALL, rot, trans, halves=radii)
    ref = oracle.InstanceScene(3, oracle.INST_BALL, rot, trans, halves=radii)
    rays = W.random_rays(20000, 0xBA20 + rotate, -220.0, 220.0)
    got = tl.raycast_nearest(dev(rays), 500.0, normal=True)
    want = ref.raycast_nearest(rays, 500.0, threads=8)
    assert (want["leaf"] >= 0).sum() > 1000
    assert_nearest_equal(host(got["leaf"]), host(got["toi"]), want["leaf"], want["t"], "ball nearest")
    hit = want["leaf"] >= 0
    assert np.array_equal(host(got["normal"])[hit].view(np.uint32), want["normal"][hit].view(np.uint32))
    assert np.array_equal(host(tl.raycast_any(dev(rays), 500.0)).astype(bool), hit)
    r_all = ref.raycast_all(rays[:4000], 500.0)
    assert_csr(tl.raycast_all(dev(rays[:4000]), 500.0), r_all, "ball all hits")
    assert_csr(tl.raycast_crossings(dev(rays[:4000]), 500.0), r_all, "ball crossings")
    # overlap: the balls' exact test
    s = W.Stream(0xBA30 + rotate)
    c = s.range(-200.0, 200.0, 5000 * 3).reshape(-1, 3)
    h = s.range(1.0, 12.0, 5000 * 3).reshape(-1, 3)
    regions = np.concatenate([c - h, c + h], 1).astype(np.float32)
    r_ov = ref.overlap(regions)
    assert r_ov[1].shape[0] > 1000
    assert_csr(tl.overlap(dev(regions)), r_ov, "ball overlap", has_t=False)
    # shape casts with a ball probe
    casts = casts_of(rays, 0.75)
    cn = tl.shapecast_nearest(dev(casts), 500.0, point=True)
    wn = ref.shapecast_nearest(casts, 500.0)
    assert (wn["leaf"] >= 0).sum() > 1000
    assert_nearest_equal(host(cn["leaf"]), host(cn["toi"]), wn["leaf"], wn["t"], "ball cast nearest")
    ca = ref.shapecast_all(casts[:4000], 500.0)
    assert ca[1].shape[0] > 500
    assert_csr(tl.shapecast_all(dev(casts[:4000]), 500.0), ca, "ball cast all")
    tl.close()


def test_box_instances_answer_no_cast_and_ball_batches(ctx):
    print("case: box instance casts, batch set over ball instances", flush=True)
    rot, trans, radii = ball_scene(3000, 0xBA40, True)
    boxes = sq_amd.Tlas(ctx, 3, sq_amd.Tlas.BOX, rot, trans, halves=np.tile(radii[:, None], (1, 3)))
    rays = W.random_rays(3000, 0xBA41, -220.0, 220.0)
    casts = dev(casts_of(rays, 0.5))
    cn = boxes.shapecast_nearest(casts, 500.0)
    assert (host(cn["leaf"]) == -1).all() and np.isinf(host(cn["toi"])).all()
    off, leaf, _ = boxes.shapecast_all(casts, 500.0)
    assert int(host(off)[-1]) == 0
    balls = sq_amd.Tlas(ctx, 3, sq_amd.Tlas.BALL, rot, trans, halves=radii)
    batches = [dev(W.random_rays(n, 0xBA50 + n, -220.0, 220.0)) for n in (1, 777, 5000)]
    bs = sq_amd.BatchSet(ctx, [(balls, r, 500.0) for r in batches])
    bs.submit()
    ctx.synchronize()
    for i, r in enumerate(batches):
        w = balls.raycast_nearest(r, 500.0)
        assert np.array_equal(host(bs.leaf[i]), host(w["leaf"]))
        assert np.array_equal(host(bs.toi[i]).view(np.uint32), host(w["toi"]).view(np.uint32))
    bs.close()
    # refit with new centres and radii answers like a fresh oracle
    trans2 = trans + np.float32(3.0)
    radii2 = radii * np.float32(1.5)
    balls.refit(rot, trans2, radii2)
    ref = oracle.InstanceScene(3, oracle.INST_BALL, rot, trans2, halves=radii2)
    got = balls.shapecast_nearest(casts, 500.0)
    want = ref.shapecast_nearest(casts_of(rays, 0.5), 500.0)
    assert_nearest_equal(host(got["leaf"]), host(got["toi"]), want["leaf"], want["t"], "ball cast after refit")


def test_ball_instance_errors_and_empty(ctx):
    print("case: ball instance refusals and an empty TLAS", flush=True)
    rot2 = np.tile(np.eye(2, dtype=np.float32), (4, 1, 1))
    with pytest.raises(sq_amd.SqError):
        sq_amd.Tlas(ctx, 2, sq_amd.Tlas.BALL, rot2, np.zeros((4, 2), np.float32), halves=np.ones(4, np.float32))
    rot3 = np.tile(np.eye(3, dtype=np.float32), (4, 1, 1))
    with pytest.raises(sq_amd.SqError):
        sq_amd.Tlas(ctx, 3, sq_amd.Tlas.BALL, rot3, np.zeros((4, 3), np.float32),
                    halves=np.array([1, -1, 1, 1], np.float32))
    empty = sq_amd.Tlas(ctx, 3, sq_amd.Tlas.BALL, np.zeros((0, 3, 3), np.float32), np.zeros((0, 3), np.float32),
                        halves=np.zeros(0, np.float32))
    casts = dev(casts_of(W.random_rays(100, 0xBA60, -10.0, 10.0), 0.5))
    got = empty.shapecast_nearest(casts, 100.0)
    assert (host(got["leaf"]) == -1).all() and np.isinf(host(got["toi"])).all()
    off, _, _ = empty.shapecast_all(casts, 100.0)
    assert (host(off) == 0).all()
    empty.close()
