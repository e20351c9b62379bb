"""Device build (sq_bvh_build_device, sq_bvh_rebuild) against the host build.

The device build must produce the reference's tree node for node: the same
child_a / child_b / prim_count / prim_indices (tree.rs:39-49,112-163), the same
node bounds bit for bit, the same stats (quality is an order-dependent sum over
the node array, so it also pins node order and bounds), and therefore the same
query results.  The host build is itself pinned to the reference by the
oracle tests; here the cases add ties in the centroid keys (stability of the
per-level sort), coincident centroids (leaves above LEAF_SIZE, chains of
synthetic pairs over 16 primitives), trees that are one leaf, and every kind
and several dimensions.  A rebuild from moved primitives equals a fresh build
of the moved primitives and answers queries like the CPU oracle.
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


def quantized_boxes(n, seed, dim=3):
    """Boxes on a coarse grid: many equal centroid coordinates (sort ties)."""
    s = W.Stream(seed)
    c = np.floor(s.range(0.0, 8.0, n * dim)).reshape(n, dim).astype(np.float32)
    h = np.full((n, dim), 0.25, np.float32)
    return np.concatenate([c - h, c + h], 1).astype(np.float32)


def coincident(n, others=0, seed=5):
    p = np.tile(np.array([[1.0, 2.0, 3.0, 1.5, 2.5, 3.5]], np.float32), (n, 1))
    if others:
        p = np.concatenate([p, W.boxes_scene(n=others, seed=seed).prims])
    return p


def balls(n, dim, seed):
    s = W.Stream(seed)
    return np.concatenate([s.range(-40.0, 40.0, n * dim).reshape(n, dim), s.range(0.2, 3.0, n).reshape(n, 1)],
                          1).astype(np.float32)


CASES = {
    "boxes3": (3, sq_amd.AABB, lambda: W.boxes_scene(n=20000, seed=0xB1).prims),
    "boxes6": (6, sq_amd.AABB, lambda: W.boxes_scene(n=30000, dim=6, lo=-7.0, hi=7.0, seed=0xB2).prims),
    "boxes2": (2, sq_amd.AABB, lambda: W.boxes_scene(n=7000, dim=2, seed=0xB3).prims),
    "balls3": (3, sq_amd.SPHERE, lambda: balls(9000, 3, 0xB4)),
    "balls5": (5, sq_amd.SPHERE, lambda: balls(4000, 5, 0xB5)),
    "tris": (3, sq_amd.TRIANGLE, lambda: W.heightfield_scene(grid=129, extent=64.0).prims),
    "ties": (3, sq_amd.AABB, lambda: quantized_boxes(12000, 0xB6)),
    "coincident40": (3, sq_amd.AABB, lambda: coincident(40)),
    "coincident1000+500": (3, sq_amd.AABB, lambda: coincident(1000, 500)),
    "coincident17": (3, sq_amd.AABB, lambda: coincident(17)),
    "n1": (3, sq_amd.AABB, lambda: W.boxes_scene(n=1, seed=1).prims),
    "n3": (3, sq_amd.AABB, lambda: W.boxes_scene(n=3, seed=2).prims),
    "n5": (3, sq_amd.AABB, lambda: W.boxes_scene(n=5, seed=3).prims),
    "n257": (3, sq_amd.AABB, lambda: W.boxes_scene(n=257, seed=4).prims),
}


def assert_same_tree(a, b, what):
    ea, eb = a.export(), b.export()
    names = ["bounds", "child_a", "child_b", "prim_count", "prim_indices"]
    for name, x, y in zip(names, ea, eb):
        assert x.shape == y.shape, f"{what}: {name} shape {x.shape} vs {y.shape}"
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), f"{what}: {name} differs"
    sa, sb = a.stats(), b.stats()
    for k in ("node_count", "leaf_count", "prim_count", "max_depth", "bytes"):
        assert sa[k] == sb[k], f"{what}: stats {k} {sa[k]} vs {sb[k]}"
    assert np.float32(sa["quality"]).view(np.uint32) == np.float32(sb["quality"]).view(np.uint32), f"{what}: quality"


@pytest.mark.parametrize("name", list(CASES))
def test_device_build_equals_host_build(ctx, name):
    dim, kind, make = CASES[name]
    prims = make()
    print(f"case: device build {name} ({prims.shape[0]} prims, D={dim})", flush=True)
    hb = sq_amd.DeviceBvh.build(ctx, dim, kind, prims)
    db = sq_amd.DeviceBvh.build_device(ctx, dim, kind, dev(prims))
    assert_same_tree(hb, db, name)
    # the host reference tree (no GPU) too
    t = sq_amd.Tree.build(dim, kind, prims)
    te, de = t.export(), db.export()
    for x, y in zip(te[1:], de[1:]):
        assert np.array_equal(x, y), f"{name}: host Tree topology differs"
    t.close()
    half = 12.0 if dim >= 5 else 60.0
    rays = W.random_rays(4000, 0xBE + dim, -half, half, dim=dim)
    rh = hb.raycast_nearest(dev(rays), 500.0)
    rd = db.raycast_nearest(dev(rays), 500.0)
    assert np.array_equal(host(rh["leaf"]), host(rd["leaf"]))
    assert np.array_equal(host(rh["toi"]).view(np.uint32), host(rd["toi"]).view(np.uint32))


def test_rebuild_moved_prims_equals_fresh_build_and_oracle(ctx):
    print("case: rebuild of 20k moving boxes over 3 frames", flush=True)
    base = W.boxes_scene(n=20000, seed=0xB7).prims
    vel = W.Stream(0xB8).range(-20.0, 20.0, 20000 * 3).reshape(20000, 3).astype(np.float32)
    db = sq_amd.DeviceBvh.build_device(ctx, 3, sq_amd.AABB, dev(base))
    rays = W.random_rays(6000, 0xB9, -150.0, 150.0)
    for frame in (1, 2, 3):
        moved = base.copy()
        moved[:, :3] += vel * frame
        moved[:, 3:] += vel * frame
        db.rebuild(dev(moved))
        fresh = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.AABB, moved)
        assert_same_tree(fresh, db, f"rebuild frame {frame}")
        r = db.raycast_nearest(dev(rays), 400.0)
        ref = oracle.Scene(3, oracle.BOX, moved).raycast_nearest(rays, 400.0, threads=8)
        assert (ref["leaf"] >= 0).sum() > 300
        assert_nearest_equal(host(r["leaf"]), host(r["toi"]), ref["leaf"], ref["t"], f"rebuild frame {frame}")


def test_device_build_c1_heightfield(ctx):
    """The 2,097,152-triangle heightfield: the device tree equals the host tree."""
    import time
    import torch
    print("case: device build of the 2M-triangle heightfield", flush=True)
    sc = W.heightfield_scene()
    prims = dev(sc.prims)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    db = sq_amd.DeviceBvh.build_device(ctx, 3, sq_amd.TRIANGLE, prims)
    t_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    hb = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, sc.prims)
    t_host = time.perf_counter() - t0
    print(f"device build {t_dev * 1e3:.1f} ms, host build + upload {t_host * 1e3:.1f} ms", flush=True)
    assert_same_tree(hb, db, "C1 heightfield")
    rays = dev(W.camera_rays(sc, res=256, seed=7))
    a, b = hb.raycast_nearest(rays, 2000.0), db.raycast_nearest(rays, 2000.0)
    assert np.array_equal(host(a["leaf"]), host(b["leaf"]))
    assert np.array_equal(host(a["toi"]).view(np.uint32), host(b["toi"]).view(np.uint32))


def test_batch_set_survives_rebuild_and_refuses_moved_scene(ctx):
    """A batch set over a device-built tree answers correctly after rebuilds
    (the tree keeps its device arrays); over a host-built tree, whose arrays
    move on the first rebuild, submit refuses instead of walking stale memory."""
    print("case: batch set over rebuilt trees", flush=True)
    base = W.boxes_scene(n=30000, seed=0xBA).prims
    vel = W.Stream(0xBB).range(-10.0, 10.0, 30000 * 3).reshape(30000, 3).astype(np.float32)
    db = sq_amd.DeviceBvh.build_device(ctx, 3, sq_amd.AABB, dev(base))
    rays = [dev(W.random_rays(n, 0xBC + n, -150.0, 150.0)) for n in (5000, 777, 20000)]
    bs = sq_amd.BatchSet(ctx, [(db, r, 400.0) for r in rays])
    for frame in (1, 2):
        moved = base.copy()
        moved[:, :3] += vel * frame
        moved[:, 3:] += vel * frame
        db.rebuild(dev(moved))
        bs.submit()
        ctx.synchronize()
        for i, r in enumerate(rays):
            w = db.raycast_nearest(r, 400.0)
            assert np.array_equal(host(bs.leaf[i]), host(w["leaf"])), f"frame {frame} batch {i}"
            assert np.array_equal(host(bs.toi[i]).view(np.uint32), host(w["toi"]).view(np.uint32))
    bs.close()
    hb = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.AABB, base)
    bs2 = sq_amd.BatchSet(ctx, [(hb, rays[0], 400.0)])
    hb.rebuild(dev(base))
    with pytest.raises(sq_amd.SqError):
        bs2.submit()
    bs2.close()


def test_wide_walk_over_device_and_rebuilt_trees(ctx):
    """The 4-wide walk (tuning `wide`) over a device-built tree, before and
    after a rebuild and a refit, equals the pair walk over the host build."""
    print("case: 4-wide walk over device-built trees", flush=True)
    base = W.boxes_scene(n=30000, seed=0xC1).prims
    db = sq_amd.DeviceBvh.build_device(ctx, 3, sq_amd.AABB, dev(base))
    rays = dev(W.random_rays(20000, 0xC2, -150.0, 150.0))
    saved = ctx.tuning()
    try:
        for step in range(3):
            prims = base + np.float32(step * 2.5)
            if step == 1:
                db.rebuild(dev(prims))
            elif step == 2:
                db.refit(dev(prims))
            ref = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.AABB, prims)
            ctx.set_tuning(**dict(saved, wide=0))
            want = ref.raycast_nearest(rays, 400.0)
            ctx.set_tuning(**dict(saved, wide=1))
            got = db.raycast_nearest(rays, 400.0)
            anyh = db.raycast_any(rays, 400.0)
            assert np.array_equal(host(got["leaf"]), host(want["leaf"])), f"step {step}"
            assert np.array_equal(host(got["toi"]).view(np.uint32), host(want["toi"]).view(np.uint32)), f"step {step}"
            assert np.array_equal(host(anyh).astype(bool), host(want["leaf"]) >= 0)
    finally:
        ctx.set_tuning(**saved)
