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
    o = oracle.InstanceScene(dim, oracle.INST_BOX, rot, tr, halves=half)
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


def identity_boxes(dim, n, seed, span):
    """Axis-aligned box instances with non-dyadic centres and half extents:
    every instance face lies on its world box's face, where the local and the
    world slab round differently."""
    s = W.Stream(seed)
    tr = s.range(-span, span, n * dim).reshape(n, dim)
    half = s.range(0.3, 2.7, n * dim).reshape(n, dim)
    rot = np.broadcast_to(np.eye(dim, dtype=np.float32), (n, dim, dim)).copy()
    return rot, tr, half


@pytest.mark.parametrize("dim,scene", [(3, "identity"), (2, "identity"), (3, "plane")])
def test_box_instances_reference_and_guarded_modes(ctx, dim, scene):
    """Box instances whose local hit lies an ulp outside the world box's slab
    interval (faces on the world box, or the plane rotations of two_level.rs):
    the default mode returns the reference Tlas's answer (unguarded instance
    test, the reference's order), bit for bit, on every ray; SQ_HITS_GUARDED
    returns the guarded answer; the two differ on some rays of these scenes."""
    print(f"case: box instances D={dim}, {scene}, both hit modes", flush=True)
    if scene == "identity":
        rot, tr, half = identity_boxes(dim, 2000, 0x1DE0 + dim, SPAN[dim])
    else:
        rot, tr, half = box_instances(dim, 2000, 0x91A0 + dim)
    rays = W.random_rays(20000, 0x7B + dim, -SPAN[dim] - 20.0, SPAN[dim] + 20.0, dim=dim)
    t = sq_amd.Tlas(ctx, dim, sq_amd.Tlas.BOX, rot, tr, halves=half)
    plain = oracle.InstanceScene(dim, oracle.INST_BOX, rot, tr, halves=half).raycast_nearest(rays, 500.0, threads=8)
    guarded = oracle.InstanceScene(dim, oracle.INST_BOX, rot, tr, halves=half, guarded=True).raycast_nearest(
        rays, 500.0, threads=8)
    got = t.raycast_nearest(dev(rays), 500.0, normal=True)
    assert_nearest_equal(host(got["leaf"]), host(got["toi"]), plain["leaf"], plain["t"], f"{scene} boxes, reference mode")
    hit = plain["leaf"] >= 0
    assert same_bits(host(got["normal"])[hit], plain["normal"][hit]).all()
    assert np.array_equal(host(t.raycast_any(dev(rays), 500.0)).astype(bool), hit)
    try:
        ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED)
        g = t.raycast_nearest(dev(rays), 500.0)
        assert_nearest_equal(host(g["leaf"]), host(g["toi"]), guarded["leaf"], guarded["t"], f"{scene} boxes, guarded mode")
    finally:
        ctx.set_hit_mode(sq_amd.SQ_HITS_REFERENCE)
    differ = int((plain["leaf"] != guarded["leaf"]).sum())
    print(f"rays whose reference answer the guard changes: {differ} of {len(rays)}", flush=True)
    assert differ > 0
    t.close()


def test_mesh_instances_guarded_mode(ctx):
    """SQ_HITS_GUARDED over mesh instances equals the oracle's guarded set
    (guarded triangles, instance hits kept inside the world boxes)."""
    print("case: mesh instances, guarded mode", flush=True)
    sc = W.instanced_scene(n_meshes=4, n_inst=300, tri_lo=200, tri_hi=2000, extent=80.0)
    meshes = [sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, m) for m in sc["meshes"]]
    t = sq_amd.Tlas(ctx, 3, sq_amd.Tlas.MESH, sc["rot"], sc["trans"], meshes=meshes, mesh_of=sc["mesh_of"], scale=sc["scale"])
    o = oracle.InstanceScene(3, oracle.INST_MESH, sc["rot"], sc["trans"], mesh_tris=sc["meshes"], mesh_of=sc["mesh_of"],
                             scale=sc["scale"], guarded=True)
    rays = W.random_rays(30000, 0xC3, -100.0, 100.0)
    ref = o.raycast_nearest(rays, 1000.0, threads=8)
    assert (ref["leaf"] >= 0).sum() > 500
    try:
        ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED)
        got = t.raycast_nearest(dev(rays), 1000.0, sub=True)
        assert_nearest_equal(host(got["leaf"]), host(got["toi"]), ref["leaf"], ref["t"], "tlas mesh, guarded mode")
        hit = ref["leaf"] >= 0
        assert np.array_equal(host(got["sub"])[hit], ref["sub"][hit])
    finally:
        ctx.set_hit_mode(sq_amd.SQ_HITS_REFERENCE)
    t.close()
