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


def test_c2_full_scene_frames(ctx):
    """The instanced scene (16 meshes, 4,096 instances) and 100,000 moving
    boxes; three frames of refit + two batch sets recorded in a graph; each
    frame's batches against the oracle on the moved boxes / the instances."""
    import torch
    print("case: C2 scene, 3 graph-replayed frames of refit + batches", flush=True)
    sc = W.instanced_scene()
    s = W.Stream(0xC2C2_7E57)
    boxes = W.boxes_scene(n=100_000, seed=int(s.u64(1)[0]), lo=-1000.0, hi=1000.0, hmin=0.5, hmax=5.0).prims
    vel = s.range(-5.0, 5.0, 100_000 * 3).reshape(100_000, 3).astype(np.float32)
    sizes = np.round(np.exp(s.range(math.log(1e3), math.log(3e4), 8).astype(np.float64))).astype(np.int64)
    rays_h = [W.random_rays(int(n), int(rs), -1500.0, 1500.0) for n, rs in zip(sizes, s.u64(8))]
    # batch 0: rays aimed at instance centres, so the instance check sees many hits
    n0 = rays_h[0].shape[0]
    o0 = s.range(-1500.0, 1500.0, n0 * 3).reshape(n0, 3).astype(np.float32)
    tgt = sc["trans"][(s.u64(n0) % np.uint64(sc["trans"].shape[0])).astype(np.int64)].astype(np.float32)
    rays_h[0] = W.rays_array(o0, W.normalize_or_zero(tgt - o0))
    stream = torch.cuda.Stream()
    ctx.set_stream(stream.cuda_stream)
    try:
        meshes = [sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, m) for m in sc["meshes"]]
        tlas = sq_amd.Tlas(ctx, 3, sq_amd.Tlas.MESH, sc["rot"], sc["trans"], meshes=meshes, mesh_of=sc["mesh_of"],
                           scale=sc["scale"])
        bvh = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.AABB, boxes)
        prims_dev = dev(boxes)
        rays = [dev(r) for r in rays_h]
        bt = sq_amd.BatchSet(ctx, [(tlas, rays[i], 4000.0) for i in range(0, 8, 2)])
        bb = sq_amd.BatchSet(ctx, [(bvh, rays[i], 4000.0) for i in range(1, 8, 2)])
        bvh.refit(prims_dev); bt.submit(); bb.submit()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            bvh.refit(prims_dev)
            bt.submit()
            bb.submit()
        oi = oracle.InstanceScene(3, oracle.INST_MESH, sc["rot"], sc["trans"], mesh_tris=sc["meshes"], mesh_of=sc["mesh_of"],
                                  scale=sc["scale"])
        inst_ref = [oi.raycast_nearest(rays_h[i], 4000.0, threads=8) for i in range(0, 8, 2)]
        assert int((inst_ref[0]["leaf"] >= 0).sum()) > 300
        for f in (0, 7, 59):
            moved = boxes.copy()
            d = vel * np.float32(f)
            moved[:, :3] += d
            moved[:, 3:] += d
            prims_dev.copy_(torch.from_numpy(moved))
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            ob = oracle.Scene(3, oracle.BOX, moved)
            for j, i in enumerate(range(1, 8, 2)):
                ref = ob.raycast_nearest(rays_h[i], 4000.0, threads=8)
                assert_nearest_equal(host(bb.leaf[j]), host(bb.toi[j]), ref["leaf"], ref["t"], f"frame {f} boxes batch {i}")
            for j, ref in enumerate(inst_ref):
                assert_nearest_equal(host(bt.leaf[j]), host(bt.toi[j]), ref["leaf"], ref["t"], f"frame {f} instances batch {j}")
        bt.close(); bb.close()
    finally:
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)
