"""Variable-length queries (all hits, crossings, overlap, cast-all) through the
pooled collect: queries split into sub-walks a few levels down the tree (or
not), one persistent walk per call appending to staging blocks, then scan,
scatter and a per-query sort.

Its CSR output must equal the one-query-per-thread count/fill kernels
(rays_per_lane = 0, pinned to the CPU oracle by test_gpu_parity.py) bit for
bit under every tuning, must equal the oracle directly, must survive a
staging area too small for the result (the walk runs again with room for the
exact total), and must report the total when the caller's capacity is short.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
import sq_amd
from sq_amd import capi
from sq_amd import workloads as W

from helpers import csr_equal

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def host(t):
    return t.cpu().numpy()


def _scenes(ctx):
    box3 = W.boxes_scene(n=6000, dim=3, lo=-30.0, hi=30.0, hmin=0.5, hmax=3.0).prims
    box6 = W.boxes_scene(n=6000, dim=6, lo=-7.0, hi=7.0, hmin=0.3, hmax=2.0).prims
    s = W.Stream(0xC011)
    balls = np.concatenate([s.range(-30.0, 30.0, 4000 * 3).reshape(-1, 3), s.range(0.5, 4.0, 4000).reshape(-1, 1)],
                           1).astype(np.float32)
    return {
        "box3": (3, sq_amd.AABB, box3, sq_amd.DeviceBvh.build(ctx, 3, sq_amd.AABB, box3)),
        "box6": (6, sq_amd.AABB, box6, sq_amd.DeviceBvh.build(ctx, 6, sq_amd.AABB, box6)),
        "ball3": (3, sq_amd.SPHERE, balls, sq_amd.DeviceBvh.build(ctx, 3, sq_amd.SPHERE, balls)),
    }


def _queries(dim, half, seed):
    s = W.Stream(seed)
    rays = W.random_rays(7000, seed, -half - 8.0, half + 8.0, dim=dim)
    c = s.range(-half, half, 5000 * dim).reshape(-1, dim)
    h = s.range(0.5, half / 3.0, 5000 * dim).reshape(-1, dim)
    regions = np.concatenate([c - h, c + h], 1).astype(np.float32)
    casts = np.concatenate([rays[:3000], np.full((3000, dim), -0.5, np.float32), np.full((3000, dim), 0.5, np.float32)], 1)
    return rays, regions, casts


def _all_results(b, rays, regions, casts, kind):
    out = {}
    out["all"] = b.raycast_all(dev(rays), 300.0)
    if kind == sq_amd.AABB:
        out["cross"] = b.raycast_all(dev(rays), 300.0, crossings=True)
    out["overlap"] = b.overlap(dev(regions))
    out["cast_all"] = b.shapecast_all(dev(casts), 40.0)
    return {k: tuple(None if x is None else host(x).copy() for x in v) for k, v in out.items()}


TUNINGS = [dict(block=128, rays_per_lane=16, lds_stack=8), dict(block=64, rays_per_lane=1, lds_stack=2),
           dict(block=256, rays_per_lane=4, lds_stack=4, descend_min=0), dict(block=128, rays_per_lane=64, lds_stack=32),
           dict(block=64, rays_per_lane=8, lds_stack=2, descend_min=64), dict(block=128, rays_per_lane=16, split_depth=-1),
           dict(block=128, rays_per_lane=16, split_depth=1), dict(block=256, rays_per_lane=2, lds_stack=2, split_depth=8)]


def test_pooled_collect_equals_per_thread_kernels(ctx):
    scenes = _scenes(ctx)
    saved = ctx.tuning()
    try:
        for name, (dim, kind, prims, b) in scenes.items():
            half = 7.0 if dim == 6 else 30.0
            rays, regions, casts = _queries(dim, half, 0xC0 + dim)
            print(f"case: collect {name}, per-thread kernels", flush=True)
            ctx.set_tuning(block=128, rays_per_lane=0, lds_stack=8)
            want = _all_results(b, rays, regions, casts, kind)
            assert want["overlap"][0][-1] > 1000 and want["all"][0][-1] > (100 if dim == 6 else 1000)
            for t in TUNINGS:
                print(f"case: collect {name}, pooled {t}", flush=True)
                ctx.set_tuning(**t)
                got = _all_results(b, rays, regions, casts, kind)
                for q, w in want.items():
                    g = got[q]
                    assert np.array_equal(g[0], w[0]), f"{name} {q} {t}: offsets differ"
                    assert np.array_equal(g[1], w[1]), f"{name} {q} {t}: ids differ"
                    if len(w) > 2 and w[2] is not None:
                        assert np.array_equal(g[2].view(np.uint32), w[2].view(np.uint32)), f"{name} {q} {t}: toi differs"
    finally:
        ctx.set_tuning(**saved)


def test_pooled_collect_matches_oracle(ctx):
    print("case: pooled collect vs CPU oracle (D=6 boxes, D=3 balls)", flush=True)
    for dim, kind, okind in ((6, sq_amd.AABB, oracle.BOX), (3, sq_amd.SPHERE, oracle.SPHERE)):
        half = 7.0 if dim == 6 else 30.0
        if kind == sq_amd.AABB:
            prims = W.boxes_scene(n=6000, dim=6, lo=-7.0, hi=7.0, hmin=0.3, hmax=2.0).prims
        else:
            s = W.Stream(0xC012)
            prims = np.concatenate([s.range(-30.0, 30.0, 4000 * 3).reshape(-1, 3), s.range(0.5, 4.0, 4000).reshape(-1, 1)],
                                   1).astype(np.float32)
        rays, regions, casts = _queries(dim, half, 0xD0 + dim)
        b = sq_amd.DeviceBvh.build(ctx, dim, kind, prims)
        o = oracle.Scene(dim, okind, prims)
        off, leaf, toi = b.raycast_all(dev(rays), 300.0)
        roff, rleaf, rt = o.raycast_all(rays, 300.0)
        csr_equal(host(off), host(leaf), host(toi), roff, rleaf, rt, f"all D={dim}")
        off, leaf = b.overlap(dev(regions))
        roff, rleaf = o.overlap(regions)
        assert len(rleaf) > 1000
        csr_equal(host(off), host(leaf), None, roff, rleaf, None, f"overlap D={dim}")
        off, leaf, toi = b.shapecast_all(dev(casts), 40.0)
        roff, rleaf, rt = o.shapecast_all(casts, 40.0)
        csr_equal(host(off), host(leaf), host(toi), roff, rleaf, rt, f"cast-all D={dim}")


def test_staging_overflow_walks_again():
    """A fresh context's first call sizes its staging area for 16M entries;
    10,500 regions, 9,000 of them covering all 2,000 boxes, need 18M, so the
    walk runs a second time with room for the total and the lists still come
    out whole."""
    import torch
    print("case: staging overflow (18M overlap entries, fresh context)", flush=True)
    c = sq_amd.Context(0)
    try:
        prims = W.boxes_scene(n=2000, dim=3, lo=-10.0, hi=10.0).prims
        b = sq_amd.DeviceBvh.build(c, 3, sq_amd.AABB, prims)
        n = 10500
        big = np.tile(np.array([[-1e3, -1e3, -1e3, 1e3, 1e3, 1e3]], np.float32), (n, 1))
        big[::7, :] = np.array([5.0, 5.0, 5.0, 6.0, 6.0, 6.0], np.float32)        # some small regions in between
        off, leaf = b.overlap(dev(big), capacity=n * 2000)
        off = host(off)
        leaf = host(leaf)
        small = oracle.Scene(3, oracle.BOX, prims).overlap(np.array([[5, 5, 5, 6, 6, 6]], np.float32))[1]
        counts = np.diff(off)
        full = np.ones(n, bool)
        full[::7] = False
        assert (counts[full] == 2000).all()
        assert (counts[~full] == len(small)).all()
        assert int(off[-1]) > (1 << 24)
        for i in (0, 1, 7, 8, n - 1):
            seg = leaf[off[i]:off[i + 1]]
            assert np.array_equal(seg, np.arange(2000) if full[i] else small), f"query {i}"
        torch.cuda.synchronize()
    finally:
        c.close()


def test_subwalk_and_staging_overflow_walk_again():
    """split_depth 10 on a 512-leaf tree: each of 34,285 all-covering regions
    splits into 512 sub-walks, 17.6M in all, over the 13.3M the item blocks
    first hold, so the walk runs again with room for them; that walk's 68.6M
    entries then overflow the 16M staging area and it runs a third time."""
    import torch
    print("case: sub-walk and staging overflow (17.6M sub-walks, 68.6M entries, fresh context)", flush=True)
    c = sq_amd.Context(0)
    try:
        c.set_tuning(split_depth=10)
        prims = W.boxes_scene(n=2000, dim=3, lo=-10.0, hi=10.0).prims
        b = sq_amd.DeviceBvh.build(c, 3, sq_amd.AABB, prims)
        n = 40000
        big = np.tile(np.array([[-1e3, -1e3, -1e3, 1e3, 1e3, 1e3]], np.float32), (n, 1))
        big[::7, :] = np.array([5.0, 5.0, 5.0, 6.0, 6.0, 6.0], np.float32)
        off, leaf = b.overlap(dev(big), capacity=n * 2000)
        off, leaf = host(off), host(leaf)
        small = oracle.Scene(3, oracle.BOX, prims).overlap(np.array([[5, 5, 5, 6, 6, 6]], np.float32))[1]
        full = np.ones(n, bool)
        full[::7] = False
        counts = np.diff(off)
        assert (counts[full] == 2000).all() and (counts[~full] == len(small)).all()
        for i in (0, 1, 7, 13, 14, n - 2, n - 1):
            seg = leaf[off[i]:off[i + 1]]
            assert np.array_equal(seg, np.arange(2000) if full[i] else small), f"query {i}"
        torch.cuda.synchronize()
    finally:
        c.close()


def test_capacity_short_reports_total(ctx):
    print("case: capacity below the total (C ABI)", flush=True)
    import torch
    prims = W.boxes_scene(n=3000, dim=3, lo=-20.0, hi=20.0, hmin=0.5, hmax=3.0).prims
    b = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.AABB, prims)
    rays = dev(W.random_rays(5000, 0xCA9, -25.0, 25.0))
    off_ref, leaf_ref, toi_ref = b.raycast_all(rays, 300.0)
    tot = int(host(off_ref)[-1])
    assert tot > 100
    offsets = torch.empty(5001, dtype=torch.int64, device="cuda")
    leaf = torch.full((tot,), -7, dtype=torch.int32, device="cuda")
    toi = torch.empty(tot, dtype=torch.float32, device="cuda")
    total = C.c_uint64()
    lib = capi.lib()
    rc = lib.sq_raycast_all(b._h, C.c_void_p(rays.data_ptr()), 5000, 300.0, None, 0, C.c_void_p(offsets.data_ptr()),
                            C.c_void_p(leaf.data_ptr()), C.c_void_p(toi.data_ptr()), tot - 1, C.byref(total))
    assert rc == capi.SQ_ERR_CAPACITY and total.value == tot
    assert (host(leaf) == -7).all(), "nothing is filled when the capacity is short"
    rc = lib.sq_raycast_all(b._h, C.c_void_p(rays.data_ptr()), 5000, 300.0, None, 0, C.c_void_p(offsets.data_ptr()),
                            C.c_void_p(leaf.data_ptr()), C.c_void_p(toi.data_ptr()), tot, C.byref(total))
    assert rc == 0 and total.value == tot
    assert np.array_equal(host(leaf), host(leaf_ref)) and np.array_equal(host(offsets), host(off_ref))
    assert np.array_equal(host(toi).view(np.uint32), host(toi_ref).view(np.uint32))
