"""Two-level walk on the device vs the oracle's Tlas (bit-exact)."""
import math

import numpy as np
import pytest

import oracle
import sq_amd
from sq_amd import workloads as W

from helpers import assert_nearest_equal, csr_equal, same_bits

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def host(t):
    return t.cpu().numpy()


def plane_rot(dim, a, b, ang):
    r = np.eye(dim, dtype=np.float32)
    s, c = np.float32(math.sin(ang)), np.float32(math.cos(ang))
    r[a, a] = c; r[a, b] = -s; r[b, a] = s; r[b, b] = c
    return r


SPAN = {2: 40.0, 3: 40.0, 5: 9.0}


def box_instances(dim, n, seed):
    s = W.Stream(seed)
    tr = s.range(-SPAN[dim], SPAN[dim], n * dim).reshape(n, dim)
    rot = np.stack([plane_rot(dim, i % dim, (i + 1) % dim, float(a)) for i, a in enumerate(s.range(0.0, 6.28, n))])
    half = np.maximum(s.range(0.5, 2.5, n * dim).reshape(n, dim), np.float32(0.5))
    return rot, tr, half


@pytest.mark.parametrize("dim", [2, 3, 5])
def test_box_instances_nearest_any_all_refit(ctx, dim):
    print(f"case: box instances D={dim}", flush=True)
    rot, tr, half = box_instances(dim, 3000, 0xB0C5 + dim)
    o = oracle.InstanceScene(dim, oracle.INST_BOX_GUARDED, rot, tr, halves=half)
    t = sq_amd.Tlas(ctx, dim, sq_amd.Tlas.BOX, rot, tr, halves=half)
    rays = W.random_rays(20000, 0x7A + dim, -SPAN[dim] - 20.0, SPAN[dim] + 20.0, dim=dim)
    for frame in range(3):
        ref = o.raycast_nearest(rays, 500.0, threads=8)
        got = t.raycast_nearest(dev(rays), 500.0, point=True, normal=True, sub=True)
        assert_nearest_equal(host(got["leaf"]), host(got["toi"]), ref["leaf"], ref["t"], f"tlas box frame {frame}")
        hit = ref["leaf"] >= 0
        assert hit.sum() > 100
        assert same_bits(host(got["normal"])[hit], ref["normal"][hit]).all()
        assert same_bits(host(got["point"])[hit], ref["point"][hit]).all()
        assert np.array_equal(host(t.raycast_any(dev(rays), 500.0)).astype(bool), hit)
        off, leaf, toi = t.raycast_all(dev(rays[:3000]), 500.0)
        roff, rleaf, rt, _ = o.raycast_all(rays[:3000], 500.0)
        csr_equal(host(off), host(leaf), host(toi), roff, rleaf, rt, "tlas all")
        regions = np.concatenate([rays[:2000, :dim] - 5, rays[:2000, :dim] + 5], 1).astype(np.float32)
        off, leaf = t.overlap(dev(regions))
        roff, rleaf = o.overlap(regions)
        csr_equal(host(off), host(leaf), None, roff, rleaf, None, "tlas overlap")
        # move every instance, refit both
        tr = tr + W.Stream(frame + 1).range(-3.0, 3.0, tr.size).reshape(tr.shape)
        o.set_transforms(rot, tr); o.refit()
        t.refit(rot, tr)
        assert np.array_equal(t.download_bounds(), o.export()[0])
        assert t.needs_rebuild(2.0) == o.needs_rebuild(2.0)


def test_mesh_instances(ctx):
    print("case: mesh instances (small C2-style scene)", flush=True)
    sc = W.instanced_scene(n_meshes=5, n_inst=400, tri_lo=200, tri_hi=3000, extent=100.0)
    meshes = [sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, m) for m in sc["meshes"]]
    t = sq_amd.Tlas(ctx, 3, sq_amd.Tlas.MESH, sc["rot"], sc["trans"], meshes=meshes, mesh_of=sc["mesh_of"], scale=sc["scale"])
    o = oracle.InstanceScene(3, oracle.INST_MESH, sc["rot"], sc["trans"], mesh_tris=sc["meshes"], mesh_of=sc["mesh_of"],
                             scale=sc["scale"])
    rays = W.random_rays(30000, 0xC2, -120.0, 120.0)
    ref = o.raycast_nearest(rays, 1000.0, threads=8)
    got = t.raycast_nearest(dev(rays), 1000.0, normal=True, sub=True)
    hit = ref["leaf"] >= 0
    assert hit.sum() > 1000
    assert_nearest_equal(host(got["leaf"]), host(got["toi"]), ref["leaf"], ref["t"], "tlas mesh nearest")
    assert np.array_equal(host(got["sub"])[hit], ref["sub"][hit]), "triangle ids differ"
    assert same_bits(host(got["normal"])[hit], ref["normal"][hit]).all()
    assert np.array_equal(host(t.raycast_any(dev(rays), 1000.0)).astype(bool), hit)
    off, leaf, toi = t.raycast_all(dev(rays[:5000]), 1000.0)
    roff, rleaf, rt, _ = o.raycast_all(rays[:5000], 1000.0)
    csr_equal(host(off), host(leaf), host(toi), roff, rleaf, rt, "tlas mesh all")
    # rigid motion: refit and compare again
    rot2 = W.random_rotations(W.Stream(5), 400)
    tr2 = sc["trans"] + np.float32(7.0)
    t.refit(rot2, tr2)
    o.set_transforms(rot2, tr2); o.refit()
    ref = o.raycast_nearest(rays, 1000.0, threads=8)
    got = t.raycast_nearest(dev(rays), 1000.0)
    assert_nearest_equal(host(got["leaf"]), host(got["toi"]), ref["leaf"], ref["t"], "tlas mesh after refit")
