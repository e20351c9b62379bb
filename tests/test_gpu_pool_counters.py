"""Pool counters of the persistent launches.

The pooled single-call kernels (BVH nearest / any / cast, TLAS nearest) take
their chunks from the context's pool counters, which start zero and which
every launch leaves zero again (the last wave to retire clears them); batch
sets do the same with their own counters.  A launch that left them non-zero
would make the next one skip chunks, so here calls of many sizes (grids of 1
to every resident wave, chunk sizes 64 to 1024) and kinds follow each other
on one context, batch submits interleaved, and every answer must equal the
one-ray-per-thread kernel (rays_per_lane 0), which uses no pool counter.
"""
import numpy as np
import pytest

import sq_amd
from sq_amd import workloads as W

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def host(t):
    return t.cpu().numpy()


def test_calls_of_every_size_leave_the_counters_clear(ctx):
    print("case: interleaved pooled calls vs the per-thread kernel", flush=True)
    sc = W.instanced_scene(n_meshes=4, n_inst=512, tri_lo=500, tri_hi=3000, extent=300.0, seed=0x9C01)
    meshes = [sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, m) for m in sc["meshes"]]
    tlas = sq_amd.Tlas(ctx, 3, sq_amd.Tlas.MESH, sc["rot"], sc["trans"], meshes=meshes, mesh_of=sc["mesh_of"],
                       scale=sc["scale"])
    boxes = W.boxes_scene(n=40000, seed=0x9C02, lo=-300.0, hi=300.0).prims
    bvh = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.AABB, boxes)
    sizes = [1, 64, 65, 1000, 70000, 3, 300000, 513, 20000, 1]
    rays = [dev(W.random_rays(n, 0x9C10 + i, -350.0, 350.0)) for i, n in enumerate(sizes)]
    saved = ctx.tuning()
    try:
        ctx.set_tuning(block=128, rays_per_lane=0, lds_stack=8)
        want = [(bvh.raycast_nearest(r, 900.0), tlas.raycast_nearest(r, 900.0), bvh.raycast_any(r, 900.0)) for r in rays]
        bs = sq_amd.BatchSet(ctx, [(bvh, r, 900.0) for r in rays[:4]])
        for tune in (dict(block=64, rays_per_lane=1, lds_stack=8), dict(block=128, rays_per_lane=16, lds_stack=8),
                     dict(block=256, rays_per_lane=4, lds_stack=4), dict(block=128, rays_per_lane=16, wide=1)):
            ctx.set_tuning(**tune)
            for rep in range(2):
                for i, r in enumerate(rays):
                    got_b = bvh.raycast_nearest(r, 900.0)
                    got_t = tlas.raycast_nearest(r, 900.0)
                    got_a = bvh.raycast_any(r, 900.0)
                    wb, wt, wa = want[i]
                    what = f"{tune} rep {rep} call {i} (n={sizes[i]})"
                    assert np.array_equal(host(got_b["leaf"]), host(wb["leaf"])), what + ": bvh ids"
                    assert np.array_equal(host(got_b["toi"]).view(np.uint32), host(wb["toi"]).view(np.uint32)), what
                    assert np.array_equal(host(got_t["leaf"]), host(wt["leaf"])), what + ": tlas ids"
                    assert np.array_equal(host(got_t["toi"]).view(np.uint32), host(wt["toi"]).view(np.uint32)), what
                    assert np.array_equal(host(got_a), host(wa)), what + ": any-hit"
                    if i % 3 == 0:
                        bs.submit()
                        ctx.synchronize()
                        for k in range(4):
                            assert np.array_equal(host(bs.leaf[k]), host(want[k][0]["leaf"])), what + f": batch {k}"
        bs.close()
    finally:
        ctx.set_tuning(**saved)
