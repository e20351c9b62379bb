"""sq_raycast_nearest_noalloc (the allocation-free nearest entry point) against
sq_raycast_nearest, through the C ABI: ids, toi, points, normals and
sub-objects bit for bit, on two calls in a row (the second smaller, so the
shared scratch keeps the first call's contents), for a triangle scene deep
enough that its stack needs the larger LDS ring (C1-sized heightfield), in both
hit modes, and for D = 6 boxes with a filter and per-ray limits."""
import numpy as np
import pytest

import sq_amd
from sq_amd import workloads as W

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def both(bvh, rays, max_toi, **kw):
    import torch
    n = rays.shape[0]
    d = rays.device
    a = bvh.raycast_nearest(rays, max_toi, point=True, normal=True, sub=True, **kw)
    leaf = torch.full((n,), -7, dtype=torch.int32, device=d)
    toi = torch.full((n,), 7.0, dtype=torch.float32, device=d)
    pt = torch.full((n, bvh.dim), 7.0, dtype=torch.float32, device=d)
    nm = torch.full((n, bvh.dim), 7.0, dtype=torch.float32, device=d)
    sb = torch.full((n,), -7, dtype=torch.int32, device=d)
    bvh.raycast_nearest_noalloc_into(rays, max_toi, leaf, toi, point=pt, normal=nm, sub=sb, **kw)
    bvh.ctx.synchronize()
    b = dict(leaf=leaf, toi=toi, point=pt, normal=nm, sub=sb)
    for k in a:
        x, y = a[k].cpu().numpy(), b[k].cpu().numpy()
        assert np.array_equal(x.view(np.uint32) if x.dtype == np.float32 else x,
                              y.view(np.uint32) if y.dtype == np.float32 else y), k
    return int((a["leaf"] >= 0).sum().item())


@pytest.mark.parametrize("mode", ["reference", "guarded"])
def test_noalloc_equals_nearest_on_the_heightfield(ctx, mode):
    print(f"case: no-alloc nearest, 1025^2 heightfield, {mode} mode", flush=True)
    sc = W.heightfield_scene()
    bvh = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, sc.prims)
    rays = np.concatenate([W.camera_rays(sc, res=512), W.incoherent_rays(sc, 1 << 18)])
    try:
        ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED if mode == "guarded" else sq_amd.SQ_HITS_REFERENCE)
        assert both(bvh, dev(rays), 2000.0) > 100000
        small = np.ascontiguousarray(rays[::397])
        assert both(bvh, dev(small), 2000.0) > 100
    finally:
        ctx.set_hit_mode(sq_amd.SQ_HITS_REFERENCE)
    bvh.close()


def test_noalloc_equals_nearest_d6_boxes_filter_limits(ctx):
    print("case: no-alloc nearest, D=6 boxes, filter, per-ray limits", flush=True)
    sc = W.gmm_boxes_scene(n=200_000)
    # origins inside the mixture's dense cores, so that most rays meet a box
    rays = W.rays_array(W.gmm_points(sc, 100_000, 0x5C, sigma=2.0), W.unit_dirs(W.Stream(0x5D), 100_000, 6))
    bvh = sq_amd.DeviceBvh.build(ctx, 6, sq_amd.AABB, sc.prims)
    layers = (np.arange(sc.prims.shape[0], dtype=np.uint64) % np.uint64(3)) + np.uint64(1)
    bvh.set_layers(layers)
    filt = sq_amd.QueryFilter(layer_mask=0b10, exclude=5)
    lim = W.Stream(0x11).range(1.0, 800.0, rays.shape[0]).astype(np.float32)
    assert both(bvh, dev(rays), 1000.0, filt=filt, max_toi_per_ray=dev(lim)) > 200
    assert both(bvh, dev(rays[:3000]), 1000.0) > 20
    bvh.close()
