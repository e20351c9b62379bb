"""Batch sets (sq_batch_*): many nearest-hit batches over many scenes in one
persistent work-stealing launch, through the C ABI.

Each batch's results must equal the single-scene call on the same rays (which
the parity tests pin to the CPU oracle) bit for bit, for uneven batch sizes
(0, 1, 63, 64, 65, ... 200k), empty scenes, every tuning, TLAS scenes, and a
captured graph (device refit + submit) replayed with eager updates between
replays, each replay checked against the eager result and the oracle.
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


SIZES = [0, 1, 63, 64, 65, 511, 512, 513, 1000, 4096, 30000, 200000]


def check_set(bs, scenes_rays, single):
    for i, (scene, rays, tmax) in enumerate(scenes_rays):
        want = single(scene, rays, tmax)
        gl, gt = host(bs.leaf[i]), host(bs.toi[i])
        assert np.array_equal(gl, host(want["leaf"])), f"batch {i}: ids differ"
        assert np.array_equal(gt.view(np.uint32), host(want["toi"]).view(np.uint32)), f"batch {i}: toi differs"


@pytest.fixture(scope="module")
def meshes(ctx):
    ms = [W.noisy_sphere(f, 0x51 + f) for f in (4, 9, 16, 31)]
    hf = W.heightfield_scene(grid=129, extent=64.0).prims
    return [sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, m) for m in ms + [hf]]


def _queries(scenes, seed):
    s = W.Stream(seed)
    q = []
    for i, n in enumerate(SIZES):
        sc = scenes[int(s.u64(1)[0] % np.uint64(len(scenes)))]
        half = 70.0 if sc.n_prims > 10000 else 2.0
        rays = W.random_rays(n, seed + i, -half, half) if n else np.zeros((0, 6), np.float32)
        q.append((sc, dev(rays), 2000.0 if i % 3 else 1.5))
    return q


def test_bvh_batch_matches_single_calls(ctx, meshes):
    print("case: batch set over 5 triangle scenes, 12 uneven batches", flush=True)
    qs = _queries(meshes, 0xBA7C)
    bs = sq_amd.BatchSet(ctx, qs)
    info = bs.info()
    chunk = 64 * ctx.tuning()["rays_per_lane"]
    assert info["chunks"] == sum((r.shape[0] + chunk - 1) // chunk for _, r, _ in qs)
    for rep in range(2):
        bs.submit()
        ctx.synchronize()
        check_set(bs, qs, lambda sc, r, t: sc.raycast_nearest(r, t))
    bs.close()


def test_bvh_batch_matches_oracle(ctx, meshes):
    print("case: batch set vs CPU oracle", flush=True)
    hf = W.heightfield_scene(grid=129, extent=64.0)
    rays = W.random_rays(20000, 0x0AC1, -70.0, 70.0)
    bs = sq_amd.BatchSet(ctx, [(meshes[4], dev(rays), 2000.0), (meshes[0], dev(rays[:777]), 2000.0)])
    bs.submit()
    ctx.synchronize()
    ref = oracle.Scene(3, oracle.TRI, hf.prims).raycast_nearest(rays, 2000.0, threads=8)
    assert (ref["leaf"] >= 0).sum() > 1000
    assert_nearest_equal(host(bs.leaf[0]), host(bs.toi[0]), ref["leaf"], ref["t"], "batch heightfield")
    bs.close()


@pytest.mark.parametrize("block,rpl,stack", [(64, 1, 2), (128, 4, 4), (256, 16, 16), (128, 8, 8)])
def test_batch_tuning_invariance(ctx, meshes, block, rpl, stack):
    print(f"case: batch tuning block {block} rays/lane {rpl} stack {stack}", flush=True)
    old = ctx.tuning()
    ctx.set_tuning(block=block, rays_per_lane=rpl, lds_stack=stack)
    try:
        qs = _queries(meshes, 0x70E + block)
        bs = sq_amd.BatchSet(ctx, qs)
        bs.submit()
        ctx.synchronize()
        check_set(bs, qs, lambda sc, r, t: sc.raycast_nearest(r, t))
        bs.close()
    finally:
        ctx.set_tuning(**old)


def test_batch_d6_boxes_and_empty_scene(ctx):
    print("case: D=6 box scenes and an empty scene in one set", flush=True)
    sc = W.gmm_boxes_scene(n=20000, dim=6, k=8)
    a = sq_amd.DeviceBvh.build(ctx, 6, sq_amd.AABB, sc.prims)
    b = sq_amd.DeviceBvh.build(ctx, 6, sq_amd.AABB, sc.prims[:333])
    e = sq_amd.DeviceBvh.build(ctx, 6, sq_amd.AABB, np.zeros((0, 12), np.float32))
    r1 = W.random_rays(50000, 0x6D1, -500.0, 500.0, dim=6)
    r2 = W.random_rays(999, 0x6D2, -500.0, 500.0, dim=6)
    qs = [(a, dev(r1), 5000.0), (e, dev(r2), 5000.0), (b, dev(r2), 5000.0), (a, dev(r2[:5]), 100.0)]
    bs = sq_amd.BatchSet(ctx, qs)
    bs.submit()
    ctx.synchronize()
    check_set(bs, qs, lambda s, r, t: s.raycast_nearest(r, t))
    assert (host(bs.leaf[1]) == -1).all() and np.isinf(host(bs.toi[1])).all()
    ref = oracle.Scene(6, oracle.BOX, sc.prims).raycast_nearest(r1, 5000.0, threads=8)
    assert_nearest_equal(host(bs.leaf[0]), host(bs.toi[0]), ref["leaf"], ref["t"], "batch D=6")
    bs.close()


def test_tlas_batch_matches_single_calls(ctx):
    print("case: batch set over two instanced-mesh scenes", flush=True)
    sc = W.instanced_scene(n_meshes=5, n_inst=400, tri_lo=200, tri_hi=3000, extent=100.0)
    meshes = [sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, m) for m in sc["meshes"]]
    t1 = sq_amd.Tlas(ctx, 3, sq_amd.Tlas.MESH, sc["rot"], sc["trans"], meshes=meshes, mesh_of=sc["mesh_of"], scale=sc["scale"])
    t2 = sq_amd.Tlas(ctx, 3, sq_amd.Tlas.MESH, sc["rot"][:50], sc["trans"][:50], meshes=meshes, mesh_of=sc["mesh_of"][:50],
                     scale=sc["scale"][:50])
    qs = []
    for i, n in enumerate([1, 64, 65, 5000, 30000, 0, 777]):
        rays = W.random_rays(n, 0x7A5 + i, -120.0, 120.0) if n else np.zeros((0, 6), np.float32)
        qs.append((t1 if i % 2 == 0 else t2, dev(rays), 1000.0))
    bs = sq_amd.BatchSet(ctx, qs)
    bs.submit()
    ctx.synchronize()
    check_set(bs, qs, lambda s, r, t: s.raycast_nearest(r, t))
    o = oracle.InstanceScene(3, oracle.INST_MESH, sc["rot"], sc["trans"], mesh_tris=sc["meshes"], mesh_of=sc["mesh_of"],
                             scale=sc["scale"])
    rays4 = W.random_rays(30000, 0x7A5 + 4, -120.0, 120.0)
    ref = o.raycast_nearest(rays4, 1000.0, threads=8)
    assert (ref["leaf"] >= 0).sum() > 1000
    assert_nearest_equal(host(bs.leaf[4]), host(bs.toi[4]), ref["leaf"], ref["t"], "tlas batch")
    bs.close()


def test_graph_of_refit_and_batches(ctx):
    """A frame (device refit of moving boxes + a batch set over them) recorded
    once into a graph, replayed four times with the boxes moved eagerly between
    replays; every replay equals the eager query on the refit tree and the
    oracle on the moved boxes."""
    import torch
    print("case: graph capture of refit + batch submit", flush=True)
    base = W.boxes_scene(n=20000, seed=0x6A7)
    prims0 = base.prims.copy()
    vel = W.Stream(0x6A8).range(-5.0, 5.0, 20000 * 3).reshape(20000, 3).astype(np.float32)
    bvh = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.AABB, prims0)
    prims_dev = dev(prims0)
    rays = [W.random_rays(n, 0x6A9 + n, -150.0, 150.0) for n in (5000, 65536, 1234)]
    qs = [(bvh, dev(r), 400.0) for r in rays]
    stream = torch.cuda.Stream()
    ctx.set_stream(stream.cuda_stream)
    try:
        bs = sq_amd.BatchSet(ctx, qs)
        # warm-up outside capture sizes every buffer
        bvh.refit(prims_dev)
        bs.submit()
        ctx.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            bvh.refit(prims_dev)
            bs.submit()
        for frame in range(1, 5):
            moved = prims0.copy()
            d = vel * np.float32(frame)
            moved[:, :3] += d
            moved[:, 3:] += d
            prims_dev.copy_(torch.from_numpy(moved))          # eager work between replays
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            got = [(host(bs.leaf[i]).copy(), host(bs.toi[i]).copy()) for i in range(len(qs))]
            for i, (scene, r, tmax) in enumerate(qs):        # eager query on the tree the graph refit
                w = scene.raycast_nearest(r, tmax)
                ctx.synchronize()
                assert np.array_equal(got[i][0], host(w["leaf"])), f"frame {frame} batch {i}"
                assert np.array_equal(got[i][1].view(np.uint32), host(w["toi"]).view(np.uint32))
            ref = oracle.Scene(3, oracle.BOX, moved).raycast_nearest(rays[1], 400.0, threads=8)
            assert_nearest_equal(got[1][0], got[1][1], ref["leaf"], ref["t"], f"graph frame {frame}")
        bs.close()
    finally:
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)
