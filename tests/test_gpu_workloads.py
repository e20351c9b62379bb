"""The batched / instanced workload and the D=6 workload at their full scene
sizes, against the CPU oracle on query samples (the drivers in
sq_amd.benches time them and check batches against single-scene calls).
"""
import math

import numpy as np
import pytest

import oracle
import sq_amd
from sq_amd import workloads as W

from helpers import assert_nearest_equal, csr_equal

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def host(t):
    return t.cpu().numpy()


def test_c4_full_scene_samples(ctx):
    print("case: D=6, 1,000,000 mixture boxes; sampled rays, casts, overlaps", flush=True)
    sc = W.gmm_boxes_scene()
    rays, regions, casts, lim = W.c4_queries(sc, n_rays=60000, n_overlap=4000, n_cast=60000, seed=0x7E57)
    b = sq_amd.DeviceBvh.build(ctx, 6, sq_amd.AABB, sc.prims)
    o = oracle.Scene(6, oracle.BOX, sc.prims)
    got = b.raycast_nearest(dev(rays), sc.max_toi)
    ref = o.raycast_nearest(rays, sc.max_toi, threads=8)
    assert (ref["leaf"] >= 0).sum() > 50
    assert_nearest_equal(host(got["leaf"]), host(got["toi"]), ref["leaf"], ref["t"], "C4 rays")
    gc = b.shapecast_nearest(dev(casts), 50.0, max_toi_per_cast=dev(lim))
    rc = o.shapecast_nearest(casts, lim)
    assert (rc["leaf"] >= 0).sum() > 1000
    assert_nearest_equal(host(gc["leaf"]), host(gc["toi"]), rc["leaf"], rc["t"], "C4 casts")
    off, leaf = b.overlap(dev(regions))
    roff, rleaf = o.overlap(regions)
    assert int(roff[-1]) > 10000
    assert np.array_equal(host(off).astype(np.uint64), roff)
    assert np.array_equal(host(leaf).astype(np.int64), rleaf)


def _moved(boxes, vel, f):
    m = boxes.copy()
    d = vel * np.float32(f)
    m[:, :3] += d
    m[:, 3:] += d
    return m


def _c2_inputs(seed, n_batches, lo, hi, instances):
    """The C2 scenes and batches: 100,000 boxes with velocities U[-5,5]^3, and
    n_batches ray batches of sizes log-uniform in [lo, hi] (origins
    U[-1500,1500]^3, uniform directions); batch 0 aims at instance centres."""
    s = W.Stream(seed)
    boxes = W.boxes_scene(n=100_000, seed=int(s.u64(1)[0]), lo=-1000.0, hi=1000.0, hmin=0.5, hmax=5.0).prims
    vel = s.range(-5.0, 5.0, 100_000 * 3).reshape(100_000, 3).astype(np.float32)
    sizes = np.round(np.exp(s.range(math.log(lo), math.log(hi), n_batches).astype(np.float64))).astype(np.int64)
    rays_h = [W.random_rays(int(n), int(rs), -1500.0, 1500.0) for n, rs in zip(sizes, s.u64(n_batches))]
    n0 = rays_h[0].shape[0]
    o0 = s.range(-1500.0, 1500.0, n0 * 3).reshape(n0, 3).astype(np.float32)
    tgt = instances["trans"][(s.u64(n0) % np.uint64(instances["trans"].shape[0])).astype(np.int64)].astype(np.float32)
    rays_h[0] = W.rays_array(o0, W.normalize_or_zero(tgt - o0))
    return boxes, vel, sizes, rays_h


def test_c2_full_scene_frames(ctx):
    """The C2 configuration: the instanced scene (16 meshes, 4,096 instances)
    and 100,000 moving boxes; 256 batches of 1k-1M rays (37M rays) per frame,
    the two batch sets' submits recorded in one graph, 60 frames in which the
    boxes move and the tree is refit, or rebuilt on the device once its surface
    sum passes 1.5x that of the tree as last built (sq_bvh_refit_auto).  On
    three frames every batch equals its single-scene call bit for bit, and a
    sample of every batch equals the CPU oracle on the moved boxes / the
    instances."""
    import torch
    print("case: C2 scene, 256 batches of 1k-1M rays, 60 graph-replayed frames with refit-or-rebuild", flush=True)
    sc = W.instanced_scene()
    nb = 256
    boxes, vel, sizes, rays_h = _c2_inputs(0xC2C2_7E57, nb, 1e3, 1e6, sc)
    assert sizes.max() > 5e5 and sizes.sum() > 2e7
    stream = torch.cuda.Stream()
    ctx.set_stream(stream.cuda_stream)
    try:
        meshes = [sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, m) for m in sc["meshes"]]
        tlas = sq_amd.Tlas(ctx, 3, sq_amd.Tlas.MESH, sc["rot"], sc["trans"], meshes=meshes, mesh_of=sc["mesh_of"],
                           scale=sc["scale"])
        prims_dev = dev(boxes)
        bvh = sq_amd.DeviceBvh.build_device(ctx, 3, sq_amd.AABB, prims_dev)
        rays = [dev(r) for r in rays_h]
        ti, bi = list(range(0, nb, 2)), list(range(1, nb, 2))
        bt = sq_amd.BatchSet(ctx, [(tlas, rays[i], 4000.0) for i in ti])

        def record():
            b = sq_amd.BatchSet(ctx, [(bvh, rays[i], 4000.0) for i in bi])
            bt.submit(); b.submit()
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=stream):
                bt.submit()
                b.submit()
            return b, g

        bb, g = record()
        oi = oracle.InstanceScene(3, oracle.INST_MESH, sc["rot"], sc["trans"], mesh_tris=sc["meshes"], mesh_of=sc["mesh_of"],
                                  scale=sc["scale"])
        pick = [np.arange(0, rays_h[i].shape[0], max(1, rays_h[i].shape[0] // 192)) for i in range(nb)]
        inst_ref = [oi.raycast_nearest(rays_h[i][pick[i]], 4000.0, threads=8) for i in ti]
        assert int((inst_ref[0]["leaf"] >= 0).sum()) > 50
        rebuilds, rerecorded, checked = 0, 0, []
        for f in range(60):
            moved = _moved(boxes, vel, f)
            prims_dev.copy_(torch.from_numpy(moved))
            if bvh.refit_auto(prims_dev, 1.5):
                rebuilds += 1
                if not bb.valid():
                    bb.close()
                    bb, g = record()
                    rerecorded += 1
            g.replay()
            torch.cuda.synchronize()
            if f not in (0, 31, 59):
                continue
            checked.append(f)
            for j, i in enumerate(ti):
                w = tlas.raycast_nearest(rays[i], 4000.0)
                torch.cuda.synchronize()          # the call ran on the context's stream
                assert torch.equal(bt.leaf[j], w["leaf"]) and torch.equal(bt.toi[j].view(torch.int32), w["toi"].view(torch.int32)), \
                    f"frame {f}: instances batch {i} differs from its single call"
                ref = inst_ref[j]
                assert_nearest_equal(host(bt.leaf[j])[pick[i]], host(bt.toi[j])[pick[i]], ref["leaf"], ref["t"],
                                     f"frame {f} instances batch {i}")
            ob = oracle.Scene(3, oracle.BOX, moved)
            for j, i in enumerate(bi):
                w = bvh.raycast_nearest(rays[i], 4000.0)
                torch.cuda.synchronize()
                assert torch.equal(bb.leaf[j], w["leaf"]) and torch.equal(bb.toi[j].view(torch.int32), w["toi"].view(torch.int32)), \
                    f"frame {f}: boxes batch {i} differs from its single call"
                ref = ob.raycast_nearest(rays_h[i][pick[i]], 4000.0, threads=8)
                assert_nearest_equal(host(bb.leaf[j])[pick[i]], host(bb.toi[j])[pick[i]], ref["leaf"], ref["t"],
                                     f"frame {f} boxes batch {i}")
        print(f"frames checked {checked}; rebuilds {rebuilds}, graphs recorded again {rerecorded}", flush=True)
        assert rebuilds >= 2
        bt.close(); bb.close()
    finally:
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)


def test_surface_sum_and_refit_auto(ctx):
    """The device surface sum equals the f64 sum of the reference tree's
    interior half surfaces (numpy over the downloaded bounds, each term in f32
    in the reference's operation order) and the oracle's f32 quality to f32
    rounding; on a tree with an over-full leaf (synthetic chain nodes) too.
    refit_auto refits while the sum stays within the factor and rebuilds to
    exactly the tree a fresh device build of the moved boxes gives."""
    import torch
    print("case: device surface sum; refit or rebuild", flush=True)

    def numpy_sum(b):
        bounds = b.download_bounds()
        _, _, _, pc, _ = b.export()
        lo, hi = bounds[:, :3], bounds[:, 3:]
        e = (hi - lo).astype(np.float32)
        terms = np.float32(0.0) + e[:, 1] * e[:, 2]
        terms = terms + e[:, 0] * e[:, 2]
        terms = terms + e[:, 0] * e[:, 1]
        return float(np.sum(terms[pc == 0].astype(np.float64)))

    s = W.Stream(0x5AF)
    boxes = W.boxes_scene(n=20_000, seed=0x5AF0, lo=-100.0, hi=100.0, hmin=0.2, hmax=2.0).prims
    dup = np.repeat(boxes[:1], 40, axis=0)                      # 41 equal boxes: an over-full leaf
    for prims in (boxes, np.concatenate([boxes, dup])):
        b = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.AABB, prims)
        got = b.surface_sum()
        assert abs(got - numpy_sum(b)) <= 1e-12 * got
        q = oracle.Scene(3, oracle.BOX, prims).tree_stats()["quality"]
        assert abs(got - q) <= 1e-4 * got
        b.close()
    vel = s.range(-2.0, 2.0, 20_000 * 3).reshape(20_000, 3).astype(np.float32)
    pd = dev(boxes)
    b = sq_amd.DeviceBvh.build_device(ctx, 3, sq_amd.AABB, pd)
    base = b.surface_sum()
    rays = W.random_rays(20000, 0x5AF1, -120.0, 120.0)
    seen = []
    for f in range(1, 40):
        moved = _moved(boxes, vel, f)
        pd.copy_(torch.from_numpy(moved))
        before = b.surface_sum()
        r = b.refit_auto(pd, 1.3)
        seen.append(r)
        if r:
            fresh = sq_amd.DeviceBvh.build_device(ctx, 3, sq_amd.AABB, pd)
            for x, y in zip(b.export(), fresh.export()):
                assert np.array_equal(x, y), f"frame {f}: the rebuilt tree differs from a fresh build"
            base = fresh.surface_sum()
            assert b.surface_sum() == base
            fresh.close()
        else:
            assert b.surface_sum() <= 1.3 * base
        got = b.raycast_nearest(dev(rays), 400.0)
        if f % 13 == 0:
            ref = oracle.Scene(3, oracle.BOX, moved).raycast_nearest(rays, 400.0, threads=8)
            assert_nearest_equal(host(got["leaf"]), host(got["toi"]), ref["leaf"], ref["t"], f"refit_auto frame {f}")
        del before
    assert any(seen) and not all(seen)
    b.close()


def test_autotuner_on_batch_sets(ctx):
    """A small autotuner sweep over two batch sets (boxes and instances): every
    candidate's results equal the first's bit for bit (the tuner checks), the
    context ends at the fastest, and that tuning reproduces the results."""
    import torch
    from sq_amd.autotune import autotune
    print("case: autotuner sweep over batch sets", flush=True)
    sc = W.instanced_scene(n_meshes=4, n_inst=300, tri_lo=200, tri_hi=3000, extent=100.0)
    entry = ctx.tuning()
    meshes = [sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, m) for m in sc["meshes"]]
    tlas = sq_amd.Tlas(ctx, 3, sq_amd.Tlas.MESH, sc["rot"], sc["trans"], meshes=meshes, mesh_of=sc["mesh_of"], scale=sc["scale"])
    boxes = W.boxes_scene(n=20_000, seed=0xA77, lo=-100.0, hi=100.0, hmin=0.2, hmax=2.0).prims
    bvh = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.AABB, boxes)
    rays = [dev(W.random_rays(n, 0xA70 + n, -120.0, 120.0)) for n in (3000, 40000, 9000, 120000, 700, 60000)]
    try:
        def prepare():
            bt = sq_amd.BatchSet(ctx, [(tlas, r, 1000.0) for r in rays[::2]])
            bb = sq_amd.BatchSet(ctx, [(bvh, r, 1000.0) for r in rays[1::2]])
            prepare.keep = (bt, bb)

            def go():
                bt.submit()
                bb.submit()
                return tuple(bt.leaf) + tuple(bt.toi) + tuple(bb.leaf) + tuple(bb.toi)
            return go

        res = autotune(ctx, prepare, space=dict(block=(64, 256), rays_per_lane=(2, 16), lds_stack=(4, 16)), reps=2)
        assert len(res["table"]) >= 8
        assert res["ms"] == min(ms for _, ms in res["table"])
        assert ctx.tuning() == res["best"]
        go = prepare()
        out = [t.clone() for t in go()]
        torch.cuda.synchronize()
        ctx.set_tuning(**entry)
        ref = [t.clone() for t in prepare()()]
        torch.cuda.synchronize()
        for x, y in zip(out, ref):
            assert torch.equal(x.view(torch.int32) if x.dtype == torch.float32 else x,
                               y.view(torch.int32) if y.dtype == torch.float32 else y)
        print(f"autotuner: best {res['best']} {res['ms']:.3f} ms (entry {res['baseline']:.3f} ms)", flush=True)
        prepare.keep = None
    finally:
        ctx.set_tuning(**entry)


def test_autotuner_picks_the_packet_walk_for_camera_rays(ctx):
    """A batch class the tuner decides on: 4M camera rays over the 2M-triangle
    heightfield walk faster as packets (batch-bounds Morton order, one
    wave-coherent walk per 64 rays: 1.18 against 1.62 ms), 1M incoherent rays
    several times slower; the tuner picks packet for the first and not for the
    second, and every candidate gives the same answers (the tuner checks).  At
    1M camera rays the two are even (0.69 against 0.67 ms: the sort costs as
    much as the packets save), so the class is sized where the choice is
    clear."""
    from sq_amd.autotune import autotune
    print("case: autotuner, packet walk per batch class", flush=True)
    sc = W.heightfield_scene()
    bvh = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, sc.prims)
    entry = ctx.tuning()
    try:
        for name, rays_h, want in (("camera", W.camera_rays(sc, res=2048), 1), ("incoherent", W.incoherent_rays(sc, 1 << 20), 0)):
            r = dev(rays_h)

            def prepare():
                def go():
                    o = bvh.raycast_nearest(r, 2000.0)
                    return (o["leaf"], o["toi"])
                return go

            res = autotune(ctx, prepare, space=dict(packet=(0, 1)), reps=3)
            print(f"{name}: {[(c['packet'], round(ms, 3)) for c, ms in res['table']]}", flush=True)
            assert res["best"]["packet"] == want, name
            ctx.set_tuning(**entry)
    finally:
        ctx.set_tuning(**entry)
    bvh.close()
