"""HIP engine vs the CPU oracle, through the C ABI (bit-exact ids and toi)."""
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


def run_nearest(ctx, dim, kind, prims, rays, max_toi, **kw):
    b = sq_amd.DeviceBvh.build(ctx, dim, kind, prims)
    r = b.raycast_nearest(dev(rays), max_toi, point=True, normal=True, sub=True, **kw)
    ctx.synchronize()
    return b, {k: host(v) for k, v in r.items()}


def span(dim):
    """Half-width of the scene cube: smaller in high dimension so a few
    thousand primitives still meet many rays."""
    return {2: 30.0, 3: 30.0, 4: 14.0, 5: 9.0, 6: 7.0}[dim]


def ball_scene(dim, n, seed, half=50.0):
    s = W.Stream(seed)
    c = s.range(-half, half, n * dim).reshape(n, dim)
    r = s.range(0.5, 4.0, n).reshape(n, 1)
    return np.concatenate([c, r], 1).astype(np.float32)


# ---------------------------------------------------------------------------

def test_c0_boxes_nearest_any_point_normal(ctx):
    print("case: C0 10k boxes x 65,536 rays", flush=True)
    sc, rays = W.c0()
    o = oracle.Scene(3, oracle.BOX, sc.prims)
    ref = o.raycast_nearest(rays, 400.0, threads=8)
    b, got = run_nearest(ctx, 3, sq_amd.AABB, sc.prims, rays, 400.0)
    assert_nearest_equal(got["leaf"], got["toi"], ref["leaf"], ref["t"], "C0 nearest")
    hit = ref["leaf"] >= 0
    assert hit.sum() > 1000 and (~hit).sum() > 1000
    assert same_bits(got["point"][hit], ref["point"][hit]).all(), "hit points differ"
    assert same_bits(got["normal"][hit], ref["normal"][hit]).all(), "entry normals differ"
    assert (got["sub"] == -1).all()
    anyh = host(b.raycast_any(dev(rays), 400.0))
    assert np.array_equal(anyh.astype(bool), hit)


@pytest.mark.parametrize("dim", [2, 3, 4, 5, 6])
def test_balls_nearest_each_dimension(ctx, dim):
    print(f"case: balls D={dim}", flush=True)
    prims = ball_scene(dim, 4000, 0xBA11 + dim)
    rays = W.random_rays(30000, 0x5EED + dim, -60.0, 60.0, dim=dim)
    o = oracle.Scene(dim, oracle.SPHERE, prims)
    ref = o.raycast_nearest(rays, 500.0, threads=8)
    b, got = run_nearest(ctx, dim, sq_amd.SPHERE, prims, rays, 500.0)
    assert_nearest_equal(got["leaf"], got["toi"], ref["leaf"], ref["t"], f"balls D{dim}")
    hit = ref["leaf"] >= 0
    assert same_bits(got["normal"][hit], ref["normal"][hit]).all()
    assert np.array_equal(got["sub"][hit], ref["sub"][hit])
    anyh = host(b.raycast_any(dev(rays), 500.0))
    assert np.array_equal(anyh.astype(bool), hit)


def test_heightfield_small_nearest_any(ctx):
    print("case: heightfield 257^2 grid, camera + incoherent rays", flush=True)
    sc = W.heightfield_scene(grid=257, extent=128.0)
    rays = np.concatenate([W.camera_rays(sc, res=256, seed=0x11), W.incoherent_rays(sc, 65536)])
    o = oracle.Scene(3, oracle.TRI, sc.prims)
    ref = o.raycast_nearest(rays, 2000.0, threads=8)
    b, got = run_nearest(ctx, 3, sq_amd.TRIANGLE, sc.prims, rays, 2000.0)
    assert_nearest_equal(got["leaf"], got["toi"], ref["leaf"], ref["t"], "heightfield nearest")
    hit = ref["leaf"] >= 0
    assert hit.sum() > 10000
    assert same_bits(got["normal"][hit], ref["normal"][hit]).all()
    anyh = host(b.raycast_any(dev(rays), 2000.0))
    assert np.array_equal(anyh.astype(bool), hit)


def test_heightfield_full_scale_c1(ctx):
    """The 2,097,152-triangle scene with all 8M rays: 100% id and toi agreement
    with the reference walk (plain Moeller-Trumbore, the reference's order);
    SQ_HITS_GUARDED agrees with the guarded oracle on every ray too, and the
    rays where the two modes differ are counted (the guard's cost in
    exactness on this workload)."""
    print("case: C1 full heightfield, 4M coherent + 4M incoherent rays", flush=True)
    sc, coh, inc = W.c1()
    rays = np.concatenate([coh, inc])
    b = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, sc.prims)
    r = b.raycast_nearest(dev(rays), 2000.0)
    anyh = b.raycast_any(dev(rays), 2000.0)
    try:
        ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED)
        g = b.raycast_nearest(dev(rays), 2000.0)
    finally:
        ctx.set_hit_mode(sq_amd.SQ_HITS_REFERENCE)
    ctx.synchronize()
    ref = oracle.Scene(3, oracle.TRI, sc.prims).raycast_nearest(rays, 2000.0, threads=16)
    assert_nearest_equal(host(r["leaf"]), host(r["toi"]), ref["leaf"], ref["t"], "C1 nearest")
    assert np.array_equal(host(anyh).astype(bool), ref["leaf"] >= 0)
    gref = oracle.Scene(3, oracle.TRI, sc.prims, guarded=True).raycast_nearest(rays, 2000.0, threads=16)
    assert_nearest_equal(host(g["leaf"]), host(g["toi"]), gref["leaf"], gref["t"], "C1 nearest, guarded mode")
    differ = int(((gref["leaf"] != ref["leaf"]) | ~same_bits(gref["t"], ref["t"])).sum())
    print(f"C1 rays whose answer the guard changes: {differ} of {len(rays)}", flush=True)
    assert differ < 1000


@pytest.mark.parametrize("dim,seed,n,span", [(3, 0xABCD_1234, 64, 60.0), (2, 0x2222_2222, 48, 60.0),
                                             (4, 0x4444_4444, 48, 40.0)])
def test_differential_scenes_match_plain_balls(ctx, dim, seed, n, span):
    """The 200 random ball scenes per dimension of differential.rs:120-162
    (same SplitMix64 draws, the reference's own ray of each scene plus 255
    more per scene from another stream): the device's default mode equals the
    oracle's plain ball provider (the differential.rs provider,
    differential.rs:61-82) walked in the reference's order, bit for bit."""
    print(f"case: differential.rs scenes D={dim}", flush=True)
    s = W.Stream(seed)
    extra = W.Stream(seed ^ 0x5EED)
    hits = 0
    for k in range(200):
        c = s.range(-50.0, 50.0, n * dim).reshape(n, dim)
        r = s.range(0.5, 4.0, n)
        o = s.range(-span, span, dim)
        d = s.range(-1.0, 1.0, dim)
        d = W.normalize_or_zero(d[None])[0]
        prims = np.concatenate([c, r[:, None]], 1).astype(np.float32)
        more = np.concatenate([extra.range(-span, span, 255 * dim).reshape(255, dim),
                               W.normalize_or_zero(extra.range(-1.0, 1.0, 255 * dim).reshape(255, dim))], 1)
        rays = np.concatenate([np.concatenate([o, d])[None], more]).astype(np.float32)
        rays = rays[np.any(rays[:, dim:] != 0, axis=1)]
        ref = oracle.Scene(dim, oracle.SPHERE_PLAIN, prims).raycast_nearest(rays, 500.0)
        b = sq_amd.DeviceBvh.build(ctx, dim, sq_amd.SPHERE, prims)
        got = b.raycast_nearest(dev(rays), 500.0, normal=True)
        assert_nearest_equal(host(got["leaf"]), host(got["toi"]), ref["leaf"], ref["t"], f"differential D{dim} scene {k}")
        hit = ref["leaf"] >= 0
        assert same_bits(host(got["normal"])[hit], ref["normal"][hit]).all()
        hits += int(hit.sum())
        b.close()
    assert hits > (50 if dim == 4 else 1000)   # the reference's D = 4 draws are sparse


def test_hit_modes_heightfield_and_balls(ctx):
    """Both hit modes against the oracle on scenes where they differ (the
    257^2 heightfield with camera and incoherent rays; balls in D = 2):
    SQ_HITS_REFERENCE equals the plain providers walked in the reference's
    order, SQ_HITS_GUARDED the guarded providers; nearest, any-hit, and all
    hits (the same in both modes)."""
    print("case: hit modes, heightfield 257^2 and D=2 balls", flush=True)
    sc = W.heightfield_scene(grid=257, extent=128.0)
    rays_h = np.concatenate([W.camera_rays(sc, res=256, seed=0x11), W.incoherent_rays(sc, 65536)])
    s = W.Stream(0x16)
    balls2 = np.concatenate([s.range(-40.0, 40.0, 8000 * 2).reshape(-1, 2), s.range(0.3, 3.0, 8000).reshape(-1, 1)],
                            1).astype(np.float32)
    rays_b = W.random_rays(60000, 0x17, -45.0, 45.0, dim=2)
    differ = 0
    for dim, kind, okind, prims, rays, tmax in ((3, sq_amd.TRIANGLE, oracle.TRI, sc.prims, rays_h, 2000.0),
                                                (2, sq_amd.SPHERE, oracle.SPHERE, balls2, rays_b, 300.0)):
        b = sq_amd.DeviceBvh.build(ctx, dim, kind, prims)
        plain = oracle.Scene(dim, okind, prims).raycast_nearest(rays, tmax, threads=8)
        guarded = oracle.Scene(dim, okind, prims, guarded=True).raycast_nearest(rays, tmax, threads=8)
        got = b.raycast_nearest(dev(rays), tmax)
        assert_nearest_equal(host(got["leaf"]), host(got["toi"]), plain["leaf"], plain["t"], f"D{dim} reference mode")
        assert np.array_equal(host(b.raycast_any(dev(rays), tmax)).astype(bool), plain["leaf"] >= 0)
        off, leaf, toi = b.raycast_all(dev(rays[:5000]), tmax)
        roff, rleaf, rt = oracle.Scene(dim, okind, prims).raycast_all(rays[:5000], tmax)
        csr_equal(host(off), host(leaf), host(toi), roff, rleaf, rt, f"D{dim} all hits")
        try:
            ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED)
            g = b.raycast_nearest(dev(rays), tmax)
            assert_nearest_equal(host(g["leaf"]), host(g["toi"]), guarded["leaf"], guarded["t"], f"D{dim} guarded mode")
        finally:
            ctx.set_hit_mode(sq_amd.SQ_HITS_REFERENCE)
        differ += int((plain["leaf"] != guarded["leaf"]).sum())
        b.close()
    assert differ > 0


def test_wide_walk_reference_order(ctx):
    """The 4-wide walk in the default hit mode (the reference's order over
    plain narrow phases, tuning `wide`) equals the oracle's plain providers on
    scenes where hits leave their boxes (the 257^2 heightfield with camera and
    incoherent rays, balls in D = 2 and 5), nearest and any-hit, with rings
    that spill; a tree with synthetic pairs (100 coincident balls) takes the
    pair walk; at full C1 scale it equals the pair walk on all 8,388,608 rays."""
    print("case: 4-wide walk, reference order", flush=True)
    sc = W.heightfield_scene(grid=257, extent=128.0)
    rays_h = np.concatenate([W.camera_rays(sc, res=256, seed=0x11), W.incoherent_rays(sc, 65536)])
    s = W.Stream(0x16)
    balls2 = np.concatenate([s.range(-40.0, 40.0, 8000 * 2).reshape(-1, 2), s.range(0.3, 3.0, 8000).reshape(-1, 1)],
                            1).astype(np.float32)
    rays_b = W.random_rays(60000, 0x17, -45.0, 45.0, dim=2)
    balls5 = ball_scene(5, 4000, 0x55, half=10.0)
    rays_5 = W.random_rays(20000, 0x56, -15.0, 15.0, dim=5)
    dup = np.tile(np.array([[5, 0, 0, 1]], np.float32), (100, 1))
    dup[::7, 3] = 1.5
    rays_d = np.array([[0, 0, 0, 1, 0, 0], [0, 0.5, 0, 1, 0, 0], [0, 9, 0, 1, 0, 0]], np.float32)
    saved = ctx.tuning()
    cases = ((3, sq_amd.TRIANGLE, oracle.TRI, sc.prims, rays_h, 2000.0), (2, sq_amd.SPHERE, oracle.SPHERE, balls2, rays_b, 300.0),
             (5, sq_amd.SPHERE, oracle.SPHERE, balls5, rays_5, 300.0), (3, sq_amd.SPHERE, oracle.SPHERE, dup, rays_d, 100.0))
    try:
        for tuning in (dict(block=128, rays_per_lane=16, lds_stack=8, wide=1), dict(block=64, rays_per_lane=1, lds_stack=2, wide=1)):
            ctx.set_tuning(**dict(saved, **tuning))
            for dim, kind, okind, prims, rays, tmax in cases:
                b = sq_amd.DeviceBvh.build(ctx, dim, kind, prims)
                plain = oracle.Scene(dim, okind, prims).raycast_nearest(rays, tmax, threads=8)
                got = b.raycast_nearest(dev(rays), tmax, normal=True)
                assert_nearest_equal(host(got["leaf"]), host(got["toi"]), plain["leaf"], plain["t"], f"D{dim} {tuning}")
                hit = plain["leaf"] >= 0
                assert same_bits(host(got["normal"])[hit], plain["normal"][hit]).all()
                assert np.array_equal(host(b.raycast_any(dev(rays), tmax)).astype(bool), hit)
                b.close()
        sc1, coh, inc = W.c1()
        rays1 = dev(np.concatenate([coh, inc]))
        b1 = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, sc1.prims)
        ctx.set_tuning(**dict(saved, wide=1))
        wide = b1.raycast_nearest(rays1, 2000.0)
        ctx.set_tuning(**dict(saved, wide=0))
        pair = b1.raycast_nearest(rays1, 2000.0)
        assert np.array_equal(host(wide["leaf"]), host(pair["leaf"]))
        assert np.array_equal(host(wide["toi"]).view(np.uint32), host(pair["toi"]).view(np.uint32))
        b1.close()
    finally:
        ctx.set_tuning(**saved)


@pytest.mark.parametrize("wide", [0, 1])
def test_lookahead_walk(ctx, wide):
    """The look-ahead form of the reference-order walk (tuning `lookahead`)
    over the pair nodes and the 4-wide records equals the oracle's plain
    providers on scenes where hits leave their boxes (the 257^2 heightfield,
    balls in D = 2 and 5, coincident balls walked as synthetic pairs), with a
    filter and with rings that spill; at full C1 scale it equals the walk
    without look-ahead on all 8,388,608 rays."""
    print(f"case: look-ahead walk, wide {wide}", flush=True)
    sc = W.heightfield_scene(grid=257, extent=128.0)
    rays_h = np.concatenate([W.camera_rays(sc, res=256, seed=0x11), W.incoherent_rays(sc, 65536)])
    s = W.Stream(0x16)
    balls2 = np.concatenate([s.range(-40.0, 40.0, 8000 * 2).reshape(-1, 2), s.range(0.3, 3.0, 8000).reshape(-1, 1)],
                            1).astype(np.float32)
    rays_b = W.random_rays(60000, 0x17, -45.0, 45.0, dim=2)
    balls5 = ball_scene(5, 4000, 0x55, half=10.0)
    rays_5 = W.random_rays(20000, 0x56, -15.0, 15.0, dim=5)
    dup = np.concatenate([np.tile(np.array([[5, 0, 0, 1]], np.float32), (100, 1)),
                          np.tile(np.array([[-5, 3, 1, 2]], np.float32), (60, 1))])
    dup[::7, 3] = 1.5
    rays_d = W.random_rays(3000, 0x19, -8.0, 8.0)
    saved = ctx.tuning()
    cases = ((3, sq_amd.TRIANGLE, oracle.TRI, sc.prims, rays_h, 2000.0), (2, sq_amd.SPHERE, oracle.SPHERE, balls2, rays_b, 300.0),
             (5, sq_amd.SPHERE, oracle.SPHERE, balls5, rays_5, 300.0), (3, sq_amd.SPHERE, oracle.SPHERE, dup, rays_d, 100.0))
    try:
        for tuning in (dict(block=128, rays_per_lane=16, lds_stack=8), dict(block=64, rays_per_lane=1, lds_stack=2)):
            ctx.set_tuning(**dict(saved, wide=wide, lookahead=1, **tuning))
            for dim, kind, okind, prims, rays, tmax in cases:
                b = sq_amd.DeviceBvh.build(ctx, dim, kind, prims)
                plain = oracle.Scene(dim, okind, prims).raycast_nearest(rays, tmax, threads=8)
                got = b.raycast_nearest(dev(rays), tmax, normal=True)
                assert_nearest_equal(host(got["leaf"]), host(got["toi"]), plain["leaf"], plain["t"], f"D{dim} {tuning}")
                hit = plain["leaf"] >= 0
                assert same_bits(host(got["normal"])[hit], plain["normal"][hit]).all()
                b.close()
        # a filter (the pair walk's look-ahead form; the 4-wide walk takes no filter)
        layers = (np.uint64(1) << (np.arange(sc.prims.shape[0], dtype=np.uint64) % np.uint64(3)))
        b = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, sc.prims)
        b.set_layers(layers)
        ref = oracle.Scene(3, oracle.TRI, sc.prims, layers=layers).raycast_nearest(rays_h, 2000.0, mask=0b110, exclude=7,
                                                                                    threads=8)
        got = b.raycast_nearest(dev(rays_h), 2000.0, filt=sq_amd.QueryFilter(0b110, 7))
        assert_nearest_equal(host(got["leaf"]), host(got["toi"]), ref["leaf"], ref["t"], "filtered look-ahead")
        b.close()
        sc1, coh, inc = W.c1()
        rays1 = dev(np.concatenate([coh, inc]))
        b1 = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, sc1.prims)
        ctx.set_tuning(**dict(saved, wide=wide, lookahead=1))
        look = b1.raycast_nearest(rays1, 2000.0)
        ctx.set_tuning(**dict(saved, wide=wide, lookahead=0))
        plain1 = b1.raycast_nearest(rays1, 2000.0)
        assert np.array_equal(host(look["leaf"]), host(plain1["leaf"]))
        assert np.array_equal(host(look["toi"]).view(np.uint32), host(plain1["toi"]).view(np.uint32))
        b1.close()
    finally:
        ctx.set_tuning(**saved)


def test_per_ray_max_toi_and_filter(ctx):
    print("case: per-ray max_toi, layer mask, exclude", flush=True)
    sc, rays = W.c0(n_rays=20000)
    tmax = W.Stream(0x7A).range(1.0, 400.0, len(rays))
    layers = (np.uint64(1) << (np.arange(sc.prims.shape[0], dtype=np.uint64) % np.uint64(3)))
    b = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.AABB, sc.prims)
    b.set_layers(layers)
    o = oracle.Scene(3, oracle.BOX, sc.prims, layers=layers)
    for mask, excl in [(0b001, None), (0b110, None), ((1 << 64) - 1, 17), (0, None)]:
        f = sq_amd.QueryFilter(mask, excl)
        r = b.raycast_nearest(dev(rays), 400.0, filt=f)
        ref = o.raycast_nearest(rays, 400.0, mask=mask, exclude=-1 if excl is None else excl)
        assert_nearest_equal(host(r["leaf"]), host(r["toi"]), ref["leaf"], ref["t"], f"filter {mask:#x} {excl}")
    # per-ray limits: the oracle runs each ray at its own max_toi
    r = b.raycast_nearest(dev(rays), 400.0, max_toi_per_ray=dev(tmax))
    b.set_layers(None)
    o2 = oracle.Scene(3, oracle.BOX, sc.prims)
    want_leaf = np.empty(2000, np.int64); want_t = np.empty(2000, np.float32)
    for i in range(2000):
        rr = o2.raycast_nearest(rays[i:i + 1], float(tmax[i]))
        want_leaf[i] = rr["leaf"][0]; want_t[i] = rr["t"][0]
    assert_nearest_equal(host(r["leaf"])[:2000], host(r["toi"])[:2000], want_leaf[:2000], want_t[:2000], "per-ray max_toi")


@pytest.mark.parametrize("kind,dim", [(oracle.BOX, 3), (oracle.BOX, 6), (oracle.SPHERE, 3), (oracle.SPHERE, 5)])
def test_all_hits_crossings(ctx, kind, dim):
    print(f"case: all-hits / crossings kind={kind} D={dim}", flush=True)
    w = span(dim)
    if kind == oracle.BOX:
        prims = W.boxes_scene(n=5000, dim=dim, lo=-w, hi=w, hmin=0.5, hmax=3.0).prims
    else:
        prims = ball_scene(dim, 5000, 0xA11, half=w)
    rays = W.random_rays(8000, 0xA11 + dim, -w - 10.0, w + 10.0, dim=dim)
    o = oracle.Scene(dim, kind, prims)
    b = sq_amd.DeviceBvh.build(ctx, dim, kind, prims)
    for cross in (False, True):
        off, leaf, toi = b.raycast_all(dev(rays), 500.0, crossings=cross)
        roff, rleaf, rt = o.raycast_all(rays, 500.0, crossings=cross)
        assert len(rleaf) > 100
        csr_equal(host(off), host(leaf), host(toi), roff, rleaf, rt, f"all cross={cross}")


@pytest.mark.parametrize("kind,dim", [(oracle.BOX, 3), (oracle.BOX, 6), (oracle.SPHERE, 4), (oracle.TRI, 3)])
def test_overlap_sorted_ids(ctx, kind, dim):
    print(f"case: overlap kind={kind} D={dim}", flush=True)
    w = span(dim)
    if kind == oracle.BOX:
        prims = W.boxes_scene(n=20000, dim=dim, lo=-w, hi=w, hmin=0.2, hmax=2.0).prims
    elif kind == oracle.SPHERE:
        prims = ball_scene(dim, 20000, 0x0E, half=w)
    else:
        prims = W.heightfield_scene(grid=129, extent=64.0).prims
    s = W.Stream(0x0E + dim)
    c = s.range(-w, w, 3000 * dim).reshape(-1, dim)
    h = s.range(1.0, 8.0, 3000 * dim).reshape(-1, dim)
    if kind == oracle.TRI:
        c[:, 2] = 0.0
    regions = np.concatenate([c - h, c + h], 1).astype(np.float32)
    o = oracle.Scene(dim, kind, prims)
    b = sq_amd.DeviceBvh.build(ctx, dim, kind, prims)
    off, leaf = b.overlap(dev(regions))
    roff, rleaf = o.overlap(regions)
    assert len(rleaf) > 1000
    csr_equal(host(off), host(leaf), None, roff, rleaf, None, "overlap")
    # long lists exercise the block sort: one region covering everything
    big = np.concatenate([np.full((1, dim), -1e4, np.float32), np.full((1, dim), 1e4, np.float32)], 1)
    off, leaf = b.overlap(dev(big))
    assert np.array_equal(host(leaf), np.arange(prims.shape[0]))


@pytest.mark.parametrize("kind,dim", [(oracle.BOX, 3), (oracle.BOX, 6), (oracle.SPHERE, 3)])
def test_shapecast_nearest_all(ctx, kind, dim):
    print(f"case: shapecast kind={kind} D={dim}", flush=True)
    w = span(dim)
    if kind == oracle.BOX:
        prims = W.boxes_scene(n=8000, dim=dim, lo=-w, hi=w, hmin=0.3, hmax=2.0).prims
    else:
        prims = ball_scene(dim, 8000, 0x5C, half=w)
    rays = W.random_rays(6000, 0x5C + dim, -w - 10.0, w + 10.0, dim=dim)
    half = np.float32(0.5)
    casts = np.concatenate([rays, np.full((len(rays), dim), -half), np.full((len(rays), dim), half)], 1).astype(np.float32)
    tmax = W.Stream(0x5D).range(0.0, 50.0, len(rays))
    tmax = np.maximum(tmax, np.float32(1e-3)).astype(np.float32)
    o = oracle.Scene(dim, kind, prims)
    b = sq_amd.DeviceBvh.build(ctx, dim, kind, prims)
    r = b.shapecast_nearest(dev(casts), 50.0, max_toi_per_cast=dev(tmax), point=True)
    ref = o.shapecast_nearest(casts, tmax)
    assert (ref["leaf"] >= 0).sum() > 100
    assert_nearest_equal(host(r["leaf"]), host(r["toi"]), ref["leaf"], ref["t"], "shapecast nearest")
    hit = ref["leaf"] >= 0
    assert same_bits(host(r["point"])[hit], ref["point"][hit]).all()
    off, leaf, toi = b.shapecast_all(dev(casts), 50.0)
    roff, rleaf, rt = o.shapecast_all(casts, 50.0)
    csr_equal(host(off), host(leaf), host(toi), roff, rleaf, rt, "shapecast all")


@pytest.mark.parametrize("kind,dim", [(oracle.BOX, 3), (oracle.SPHERE, 6), (oracle.TRI, 3)])
def test_device_refit_matches_host_refit(ctx, kind, dim):
    """Bvh::refit on the device == the reference reverse pass, bit for bit; queries after it match."""
    print(f"case: refit kind={kind} D={dim}", flush=True)
    import torch
    if kind == oracle.BOX:
        prims = W.boxes_scene(n=30000, dim=dim).prims
    elif kind == oracle.SPHERE:
        prims = ball_scene(dim, 30000, 0xF1)
    else:
        prims = W.heightfield_scene(grid=129, extent=64.0).prims
    o = oracle.Scene(dim, kind, prims)
    b = sq_amd.DeviceBvh.build(ctx, dim, kind, prims)
    s = W.Stream(0xF17)
    for frame in range(3):
        if kind == oracle.SPHERE:
            moved = prims.copy(); moved[:, :dim] += s.range(-5.0, 5.0, prims.shape[0] * dim).reshape(-1, dim)
        else:
            step = s.range(-5.0, 5.0, prims.shape[0] * dim).reshape(-1, dim)
            moved = prims.copy()
            reps = moved.shape[1] // dim
            for k in range(reps):
                moved[:, k * dim:(k + 1) * dim] += step
        prims = moved.astype(np.float32)
        b.refit(torch.from_numpy(prims).cuda())
        o.set_prims(prims)
        o.refit()
        got = b.download_bounds()
        want = o.export()[0]
        assert np.array_equal(got, want), f"frame {frame}: refit bounds differ"
        rays = W.random_rays(20000, 0xF0 + frame, -150.0, 150.0, dim=dim)
        r = b.raycast_nearest(dev(rays), 1000.0)
        ref = o.raycast_nearest(rays, 1000.0, threads=8)
        assert_nearest_equal(host(r["leaf"]), host(r["toi"]), ref["leaf"], ref["t"], f"after refit {frame}")


def test_required_work_matches_oracle(ctx):
    print("case: required work counter", flush=True)
    sc = W.heightfield_scene(grid=257, extent=128.0)
    rays = W.incoherent_rays(sc, 30000)
    o = oracle.Scene(3, oracle.TRI, sc.prims)
    ref = o.raycast_nearest(rays, 2000.0, threads=8)
    tl = np.where(ref["leaf"] >= 0, ref["t"], np.float32(2000.0)).astype(np.float32)
    b = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, sc.prims)
    nodes, prims = b.required_work(dev(rays), dev(tl))
    rn, rp = o.required_work(rays, tl)
    assert np.array_equal(host(nodes).astype(np.uint64), rn)
    assert np.array_equal(host(prims).astype(np.uint64), rp)


def test_edge_cases(ctx):
    print("case: empty scene, empty batch, single leaf, coincident centroids, axis-parallel rays", flush=True)
    import torch
    # empty scene
    b = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.SPHERE, np.zeros((0, 4), np.float32))
    r = b.raycast_nearest(dev(np.array([[0, 0, 0, 1, 0, 0]], np.float32)), 100.0)
    assert host(r["leaf"])[0] == -1 and np.isinf(host(r["toi"])[0])
    # empty batch
    b = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.SPHERE, np.array([[5, 0, 0, 1]], np.float32))
    r = b.raycast_nearest(dev(np.zeros((0, 6), np.float32)), 100.0)
    assert r["leaf"].numel() == 0
    # single leaf: the reference doc example (lib.rs:49-54): toi 4, leaf 0; the away ray misses; max_toi clips
    rays = np.array([[0, 0, 0, 1, 0, 0], [0, 0, 0, -1, 0, 0]], np.float32)
    r = b.raycast_nearest(dev(rays), 100.0)
    assert host(r["leaf"]).tolist() == [0, -1] and host(r["toi"])[0] == np.float32(4.0)
    assert host(b.raycast_nearest(dev(rays[:1]), 3.0)["leaf"])[0] == -1
    # coincident centroids: a 40-primitive leaf (split into chains on the device); exact tie -> lowest leaf
    prims = np.tile(np.array([[5, 0, 0, 1]], np.float32), (40, 1))
    b = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.SPHERE, prims)
    assert b.stats()["node_count"] == 1
    r = b.raycast_nearest(dev(rays[:1]), 100.0)
    assert host(r["leaf"])[0] == 0
    off, leaf, toi = b.raycast_all(dev(rays[:1]), 100.0)
    assert host(leaf).tolist() == list(range(40))
    # axis-parallel rays through a box grid take the slab's parallel branch
    sc = W.boxes_scene(n=2000, lo=-20.0, hi=20.0)
    s = W.Stream(0xAA)
    o_ = s.range(-25.0, 25.0, 3000 * 3).reshape(-1, 3)
    axis = (s.u64(3000) % np.uint64(3)).astype(np.int64)
    sign = np.where(s.u64(3000) % np.uint64(2) == 0, 1.0, -1.0).astype(np.float32)
    d = np.zeros((3000, 3), np.float32); d[np.arange(3000), axis] = sign
    rays = np.concatenate([o_, d], 1).astype(np.float32)
    o = oracle.Scene(3, oracle.BOX, sc.prims)
    ref = o.raycast_nearest(rays, 400.0)
    b = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.AABB, sc.prims)
    r = b.raycast_nearest(dev(rays), 400.0)
    assert_nearest_equal(host(r["leaf"]), host(r["toi"]), ref["leaf"], ref["t"], "axis-parallel")


@pytest.mark.parametrize("tuning", [dict(block=64, rays_per_lane=0, lds_stack=8), dict(block=256, rays_per_lane=0, lds_stack=8),
                                    dict(block=64, rays_per_lane=1, lds_stack=2), dict(block=128, rays_per_lane=8, lds_stack=4),
                                    dict(block=256, rays_per_lane=64, lds_stack=32),
                                    dict(block=128, rays_per_lane=8, lds_stack=8, ray_order=1),
                                    dict(block=64, rays_per_lane=2, lds_stack=4, ray_order=1),
                                    dict(block=128, rays_per_lane=16, lds_stack=8, wide=1),
                                    dict(block=64, rays_per_lane=1, lds_stack=2, wide=1),
                                    dict(block=128, rays_per_lane=16, lds_stack=8, wide=0),
                                    dict(block=64, rays_per_lane=2, lds_stack=2, wide=0),
                                    dict(block=128, rays_per_lane=16, lds_stack=8, wide=2, lookahead=1),
                                    dict(block=64, rays_per_lane=2, lds_stack=2, wide=0, lookahead=1),
                                    dict(block=128, rays_per_lane=16, lds_stack=8, wide=2, lookahead=1, lookahead_pause=0),
                                    dict(block=128, rays_per_lane=4, lds_stack=4, wide=2, lookahead=1, lookahead_pause=64)])
def test_tuning_does_not_change_results(ctx, tuning):
    """Every kernel variant (one query per thread; pooled with LDS rings that
    spill constantly or never; rays walked in Morton order) gives the same bits."""
    print(f"case: tuning {tuning}", flush=True)
    sc = W.heightfield_scene(grid=129, extent=64.0)
    rays = np.concatenate([W.camera_rays(sc, res=128, seed=0x11), W.incoherent_rays(sc, 20000)])
    o = oracle.Scene(3, oracle.TRI, sc.prims)
    ref = o.raycast_nearest(rays, 2000.0, threads=8)
    b = sq_amd.DeviceBvh.build(ctx, 3, sq_amd.TRIANGLE, sc.prims)
    saved = ctx.tuning()
    ctx.set_tuning(**tuning)
    try:
        r = b.raycast_nearest(dev(rays), 2000.0, normal=True)
        anyh = b.raycast_any(dev(rays), 2000.0)
        prims6 = W.boxes_scene(n=3000, dim=6, lo=-7.0, hi=7.0, hmin=0.3, hmax=2.0).prims
        q6 = W.random_rays(3000, 9, -17.0, 17.0, dim=6)
        b6 = sq_amd.DeviceBvh.build(ctx, 6, sq_amd.AABB, prims6)
        casts = np.concatenate([q6, np.full((len(q6), 6), -0.5, np.float32), np.full((len(q6), 6), 0.5, np.float32)], 1)
        rc = b6.shapecast_nearest(dev(casts), 50.0)
    finally:
        ctx.set_tuning(**saved)
    assert_nearest_equal(host(r["leaf"]), host(r["toi"]), ref["leaf"], ref["t"], f"tuning {tuning}")
    hit = ref["leaf"] >= 0
    assert same_bits(host(r["normal"])[hit], ref["normal"][hit]).all()
    assert np.array_equal(host(anyh).astype(bool), hit)
    ref6 = oracle.Scene(6, oracle.BOX, prims6).shapecast_nearest(casts, 50.0)
    assert_nearest_equal(host(rc["leaf"]), host(rc["toi"]), ref6["leaf"], ref6["t"], "cast tuning")


@pytest.mark.parametrize("case", ["c0_boxes", "balls3", "balls6", "heightfield", "coincident", "filter"])
def test_reference_query_stats(ctx, case):
    """sq_raycast_nearest_stats / sq_raycast_all_stats: the reference's
    QueryStats per ray (nodes_visited, aabb_tests, narrow_tests, hits), equal
    to the oracle's counters on every ray, with the reference's nearest hit; in
    both hit modes (the counted walk is the reference's whatever the mode)."""
    print(f"case: reference query stats, {case}", flush=True)
    mask, exclude, layers = oracle.ALL_LAYERS, -1, None
    if case == "c0_boxes":
        sc = W.c0()[0]
        dim, kind, okind, prims, tmax = 3, sq_amd.AABB, oracle.BOX, sc.prims, 400.0
        rays = W.c0()[1][:20000]
    elif case.startswith("balls"):
        dim = int(case[-1])
        prims = ball_scene(dim, 20000, 0x57A7 + dim, half=span(dim) * 3)
        kind, okind, tmax = sq_amd.SPHERE, oracle.SPHERE, 500.0
        rays = W.random_rays(20000, 0x57A8 + dim, -span(dim) * 3.5, span(dim) * 3.5, dim=dim)
    elif case == "heightfield":
        hf = W.heightfield_scene(grid=257, extent=128.0)
        dim, kind, okind, prims, tmax = 3, sq_amd.TRIANGLE, oracle.TRI, hf.prims, 2000.0
        rays = np.concatenate([W.camera_rays(hf, res=128, seed=0x13), W.incoherent_rays(hf, 16384)])
    elif case == "coincident":
        prims = np.tile(np.array([[5, 0, 0, 1]], np.float32), (100, 1))
        prims[::7, 3] = 1.5
        prims = np.concatenate([prims, ball_scene(3, 500, 0xC01, half=40.0)])
        dim, kind, okind, tmax = 3, sq_amd.SPHERE, oracle.SPHERE, 100.0
        rays = np.concatenate([np.array([[0, 0, 0, 1, 0, 0], [0, 0.5, 0, 1, 0, 0]], np.float32),
                               W.random_rays(4000, 0xC02, -50.0, 50.0)])
    else:
        prims = ball_scene(3, 5000, 0xF17, half=40.0)
        dim, kind, okind, tmax = 3, sq_amd.SPHERE, oracle.SPHERE, 500.0
        rays = W.random_rays(8000, 0xF18, -50.0, 50.0)
        layers = (np.arange(5000, dtype=np.uint64) % np.uint64(3)) + np.uint64(1)
        mask, exclude = 0b10, 7
    o = oracle.Scene(dim, okind, prims)
    if layers is not None:
        o.set_layers(layers)
    ref = o.raycast_nearest(rays, tmax, mask=mask, exclude=exclude, threads=8, stats=True)
    ref_all = o.raycast_all_stats(rays, tmax, mask=mask, exclude=exclude)
    b = sq_amd.DeviceBvh.build(ctx, dim, kind, prims)
    filt = None
    if layers is not None:
        b.set_layers(layers)
        filt = sq_amd.QueryFilter(layer_mask=mask, exclude=exclude)
    for mode in (sq_amd.SQ_HITS_REFERENCE, sq_amd.SQ_HITS_GUARDED):
        ctx.set_hit_mode(mode)
        got = b.raycast_nearest_stats(dev(rays), tmax, filt=filt)
        sa = b.raycast_all_stats(dev(rays), tmax, filt=filt)
        ctx.synchronize()
        assert_nearest_equal(host(got["leaf"]), host(got["toi"]), ref["leaf"], ref["t"], f"{case} counted walk")
        assert np.array_equal(host(got["stats"]).astype(np.uint64), ref["stats"]), f"{case}: nearest counters"
        assert np.array_equal(host(sa).astype(np.uint64), ref_all), f"{case}: all-hits counters"
    ctx.set_hit_mode(sq_amd.SQ_HITS_REFERENCE)
    b.close()


@pytest.mark.parametrize("case", ["heightfield", "balls3", "balls5", "c0_boxes", "coincident"])
def test_packet_walk(ctx, case):
    """Tuning `packet` (one wave-coherent walk per 64 rays, scalar node and
    primitive loads, per-lane limits): the reference's answer on every ray in
    the default hit mode, the guarded oracle's in the guarded mode, for
    coherent camera rays and incoherent rays alike."""
    print(f"case: packet walk, {case}", flush=True)
    if case == "heightfield":
        hf = W.heightfield_scene(grid=257, extent=128.0)
        dim, kind, okind, prims, tmax = 3, sq_amd.TRIANGLE, oracle.TRI, hf.prims, 2000.0
        rays = np.concatenate([W.camera_rays(hf, res=256, seed=0x14), W.incoherent_rays(hf, 20000)])
    elif case.startswith("balls"):
        dim = int(case[-1])
        prims = ball_scene(dim, 20000, 0x9AC + dim, half=span(dim) * 3)
        kind, okind, tmax = sq_amd.SPHERE, oracle.SPHERE, 500.0
        rays = W.random_rays(30000, 0x9AD + dim, -span(dim) * 3.5, span(dim) * 3.5, dim=dim)
    elif case == "c0_boxes":
        sc, rays = W.c0(30000)
        dim, kind, okind, prims, tmax = 3, sq_amd.AABB, oracle.BOX, sc.prims, 400.0
    else:
        prims = np.tile(np.array([[5, 0, 0, 1]], np.float32), (100, 1))
        prims[::7, 3] = 1.5
        prims = np.concatenate([prims, ball_scene(3, 500, 0xC03, half=40.0)])
        dim, kind, okind, tmax = 3, sq_amd.SPHERE, oracle.SPHERE, 100.0
        rays = np.concatenate([np.array([[0, 0, 0, 1, 0, 0], [0, 0.5, 0, 1, 0, 0]], np.float32),
                               W.random_rays(4000, 0xC04, -50.0, 50.0)])
    b = sq_amd.DeviceBvh.build(ctx, dim, kind, prims)
    entry = ctx.tuning()
    try:
        for guarded in (False, True):
            ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED if guarded else sq_amd.SQ_HITS_REFERENCE)
            ref = oracle.Scene(dim, okind, prims, guarded=guarded).raycast_nearest(rays, tmax, threads=8)
            ctx.set_tuning(**dict(entry, packet=1))
            got = b.raycast_nearest(dev(rays), tmax, point=True, normal=True, sub=True)
            ctx.set_tuning(**entry)
            plain = b.raycast_nearest(dev(rays), tmax, point=True, normal=True, sub=True)
            ctx.synchronize()
            assert (ref["leaf"] >= 0).sum() > 100
            assert_nearest_equal(host(got["leaf"]), host(got["toi"]), ref["leaf"], ref["t"], f"packet {case} guarded={guarded}")
            for k in ("point", "normal", "sub"):
                assert same_bits(host(got[k]), host(plain[k])).all(), f"packet {case}: {k}"
    finally:
        ctx.set_tuning(**entry)
        ctx.set_hit_mode(sq_amd.SQ_HITS_REFERENCE)
    b.close()
