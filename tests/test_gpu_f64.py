"""The `f64` feature on the device (sq_bvh64_*: sq/src/maths/point.rs:14-18,
Scalar = f64) against the f64 oracle (oracle/sq_oracle64.cpp, the naive scans
of sq/src/query/naive.rs in double):

  * D = 2, 3, 4, 6 over BOX and SPHERE leaves: nearest (ids, toi, point and
    normal bit for bit), any-hit, all hits, shape casts (nearest with points,
    all) and overlap (CSR bit for bit), with and without a filter; the tree's
    stats equal the oracle's f64 tree;
  * a 262,144-ball scene with 1M rays: the oracle on a sample, and the
    size-independent properties on every ray (any == nearest hit, the first
    of a ray's all-hits list is its nearest hit, lists sorted);
  * refusals (triangles, D = 7, negative radii) and the empty tree;
  * refit (Bvh::refit, Tlas::refit, Tlas::needs_rebuild) against the oracle's
    tree built on the first positions and refit to the second;
  * consumer narrow phases in double (sq_narrow64, compiled at run time): the
    built-in ball's arithmetic equals the f64 SPHERE kind and the oracle's
    consumer kind bit for bit in both hit modes; balls and ellipsoids with
    shrunk boxes equal the oracle's reference-order walk and guarded walk;
    compile errors, refusals and the empty tree.
"""
import numpy as np
import pytest

import oracle
import sq_amd

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()


def host(t):
    return t.cpu().numpy()


def scene(dim, kind, n, seed):
    rng = np.random.default_rng(seed)
    span = {2: 60.0, 3: 30.0, 4: 12.0, 6: 6.0}[dim]
    c = rng.uniform(-span, span, (n, dim))
    if kind == sq_amd.AABB:
        prims = np.concatenate([c, c + rng.uniform(0.3, 3.0, (n, dim))], 1)
    else:
        prims = np.concatenate([c, rng.uniform(0.3, 2.5, (n, 1))], 1)
    o = rng.uniform(-span, span, (3 * n // 2, dim))
    d = rng.normal(size=o.shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], 1)
    h = rng.uniform(0.5, 3.0, (n // 2, dim))
    regions = np.concatenate([c[: n // 2] - h, c[: n // 2] + h], 1)
    ph = rng.uniform(0.1, 1.0, (rays.shape[0], 1))
    casts = np.concatenate([rays, np.repeat(-ph, dim, 1), np.repeat(ph, dim, 1)], 1)
    return prims, rays, regions, casts, span


def assert_csr(got, ref, what, has_t=True):
    assert np.array_equal(host(got[0]).astype(np.uint64), ref[0].astype(np.uint64)), f"{what}: offsets"
    assert np.array_equal(host(got[1]).astype(np.int64), ref[1]), f"{what}: ids"
    if has_t:
        assert np.array_equal(host(got[2]).view(np.uint64), ref[2].view(np.uint64)), f"{what}: toi"


@pytest.mark.parametrize("dim", [2, 3, 4, 6])
@pytest.mark.parametrize("kind", [sq_amd.AABB, sq_amd.SPHERE])
def test_f64_matches_oracle(ctx, dim, kind):
    print(f"case: f64 dim={dim} kind={kind}", flush=True)
    prims, rays, regions, casts, span = scene(dim, kind, 4000, 640 + 10 * dim + kind)
    layers = np.random.default_rng(dim).integers(0, 4, prims.shape[0]).astype(np.uint64)
    b = sq_amd.DeviceBvh64(ctx, dim, kind, prims)
    ref = oracle.Scene64(dim, kind, prims)
    max_toi = 4.0 * span
    st, rs = b.stats(), ref.tree_stats()
    for k in ("node_count", "leaf_count", "prim_count", "max_depth", "bytes"):
        assert st[k] == rs[k], k
    assert st["quality"] == float(np.float32(rs["quality"]))
    got = b.raycast_nearest(dev(rays), max_toi, point=True, normal=True)
    want = ref.raycast_nearest(rays, max_toi)
    hit = want["leaf"] >= 0
    assert hit.sum() > len(rays) // 10
    assert np.array_equal(host(got["leaf"]).astype(np.int64), want["leaf"])
    assert np.array_equal(host(got["toi"]).view(np.uint64), want["t"].view(np.uint64))
    assert np.array_equal(host(got["point"]).view(np.uint64), want["point"].view(np.uint64))
    assert np.array_equal(host(got["normal"]).view(np.uint64), want["normal"].view(np.uint64))
    assert np.array_equal(host(b.raycast_any(dev(rays), max_toi)).astype(bool), hit)
    r_all = ref.raycast_all(rays[:2000], max_toi)
    assert r_all[1].shape[0] > 200
    assert_csr(b.raycast_all(dev(rays[:2000]), max_toi), r_all, "all hits")
    r_ov = ref.overlap(regions)
    assert r_ov[1].shape[0] > 1000
    assert_csr(b.overlap(dev(regions)), (r_ov[0], r_ov[1], None), "overlap", has_t=False)
    cn = b.shapecast_nearest(dev(casts), max_toi, point=True)
    wc = ref.shapecast_nearest(casts, max_toi)
    assert (wc["leaf"] >= 0).sum() > len(casts) // 10
    assert np.array_equal(host(cn["leaf"]).astype(np.int64), wc["leaf"])
    assert np.array_equal(host(cn["toi"]).view(np.uint64), wc["t"].view(np.uint64))
    assert np.array_equal(host(cn["point"]).view(np.uint64), wc["point"].view(np.uint64))
    assert_csr(b.shapecast_all(dev(casts[:1500]), max_toi), ref.shapecast_all(casts[:1500], max_toi), "cast all")
    # filters: layers by original leaf id, a mask and an excluded id
    b.set_layers(layers)
    reff = oracle.Scene64(dim, kind, prims, layers=layers)
    filt = sq_amd.QueryFilter(layer_mask=0b10, exclude=int(want["leaf"][hit][0]))
    got = b.raycast_nearest(dev(rays), max_toi, filt=filt)
    want = reff.raycast_nearest(rays, max_toi, mask=0b10, exclude=filt.exclude)
    assert np.array_equal(host(got["leaf"]).astype(np.int64), want["leaf"])
    assert np.array_equal(host(got["toi"]).view(np.uint64), want["t"].view(np.uint64))
    assert_csr(b.raycast_all(dev(rays[:2000]), max_toi, filt=filt),
               reff.raycast_all(rays[:2000], max_toi, mask=0b10, exclude=filt.exclude), "filtered all hits")
    r_ov = reff.overlap(regions, mask=0b10, exclude=filt.exclude)
    assert_csr(b.overlap(dev(regions), filt=filt), (r_ov[0], r_ov[1], None), "filtered overlap", has_t=False)
    cn = b.shapecast_nearest(dev(casts), max_toi, filt=filt)
    wc = reff.shapecast_nearest(casts, max_toi, mask=0b10, exclude=filt.exclude)
    assert np.array_equal(host(cn["leaf"]).astype(np.int64), wc["leaf"])
    assert np.array_equal(host(cn["toi"]).view(np.uint64), wc["t"].view(np.uint64))
    # per-query limits
    lim = np.random.default_rng(dim + 100).uniform(0.1, max_toi, len(rays))
    got = b.raycast_nearest(dev(rays), max_toi, max_toi_per_ray=dev(lim))
    want = reff.raycast_nearest(rays, lim)
    assert np.array_equal(host(got["leaf"]).astype(np.int64), want["leaf"])
    assert np.array_equal(host(got["toi"]).view(np.uint64), want["t"].view(np.uint64))
    assert np.array_equal(host(b.raycast_any(dev(rays), max_toi, max_toi_per_ray=dev(lim))).astype(bool),
                          want["leaf"] >= 0)
    cn = b.shapecast_nearest(dev(casts), max_toi, max_toi_per_cast=dev(lim))
    wc = reff.shapecast_nearest(casts, lim)
    assert np.array_equal(host(cn["leaf"]).astype(np.int64), wc["leaf"])
    assert np.array_equal(host(cn["toi"]).view(np.uint64), wc["t"].view(np.uint64))
    assert_csr(b.shapecast_all(dev(casts[:1500]), max_toi, max_toi_per_cast=dev(lim[:1500])),
               reff.shapecast_all(casts[:1500], lim[:1500]), "cast all, per-cast limits")
    b.close()


def test_f64_large_ball_scene(ctx):
    print("case: f64 262144 balls, 1M rays", flush=True)
    rng = np.random.default_rng(0xF64)
    n, nr = 1 << 18, 1 << 20
    prims = np.concatenate([rng.uniform(-400, 400, (n, 3)), rng.uniform(0.5, 4.0, (n, 1))], 1)
    o = rng.uniform(-420, 420, (nr, 3))
    d = rng.normal(size=(nr, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], 1)
    b = sq_amd.DeviceBvh64(ctx, 3, sq_amd.SPHERE, prims)
    rd = dev(rays)
    got = b.raycast_nearest(rd, 2000.0)
    leaf, toi = host(got["leaf"]).astype(np.int64), host(got["toi"])
    assert (leaf >= 0).sum() > nr // 4
    # the oracle on an evenly spaced sample
    sample = np.arange(0, nr, nr // 1024)
    want = oracle.Scene64(3, sq_amd.SPHERE, prims).raycast_nearest(rays[sample], 2000.0)
    assert np.array_equal(leaf[sample], want["leaf"])
    assert np.array_equal(toi[sample].view(np.uint64), want["t"].view(np.uint64))
    # size-independent properties over every ray
    anyh = host(b.raycast_any(rd, 2000.0)).astype(bool)
    assert np.array_equal(anyh, leaf >= 0)
    off, ids, ts = b.raycast_all(rd, 2000.0)
    off, ids, ts = host(off).astype(np.int64), host(ids).astype(np.int64), host(ts)
    cnt = np.diff(off)
    assert np.array_equal(cnt > 0, leaf >= 0)
    first = off[:-1][cnt > 0]
    assert np.array_equal(ids[first], leaf[cnt > 0])
    assert np.array_equal(ts[first].view(np.uint64), toi[cnt > 0].view(np.uint64))
    same_ray = np.repeat(np.arange(nr), cnt)
    nxt = np.arange(len(ids) - 1)
    inner = same_ray[nxt] == same_ray[nxt + 1]
    a, c = nxt[inner], nxt[inner] + 1
    assert ((ts[a] < ts[c]) | ((ts[a] == ts[c]) & (ids[a] < ids[c]))).all()
    b.close()


def test_f64_refusals_and_empty(ctx):
    print("case: f64 refusals and the empty tree", flush=True)
    with pytest.raises(sq_amd.SqError):
        sq_amd.DeviceBvh64(ctx, 3, sq_amd.TRIANGLE, np.zeros((4, 9)))
    with pytest.raises(sq_amd.SqError):
        sq_amd.DeviceBvh64(ctx, 7, sq_amd.AABB, np.zeros((4, 14)))
    with pytest.raises(sq_amd.SqError):
        sq_amd.DeviceBvh64(ctx, 3, sq_amd.SPHERE, np.array([[0.0, 0.0, 0.0, -1.0]]))
    empty = sq_amd.DeviceBvh64(ctx, 3, sq_amd.AABB, np.zeros((0, 6)))
    rays = dev(np.array([[0.0, 0.0, 0.0, 1.0, 0.0, 0.0]] * 5))
    got = empty.raycast_nearest(rays, 10.0, point=True)
    assert (host(got["leaf"]) == -1).all() and np.isinf(host(got["toi"])).all()
    assert (host(got["point"]) == 0).all()
    off, _, _ = empty.raycast_all(rays, 10.0)
    assert (host(off) == 0).all()
    casts = dev(np.array([[0.0, 0.0, 0.0, 1.0, 0.0, 0.0, -0.5, -0.5, -0.5, 0.5, 0.5, 0.5]] * 5))
    got = empty.shapecast_nearest(casts, 10.0)
    assert (host(got["leaf"]) == -1).all() and np.isinf(host(got["toi"])).all()
    assert empty.stats()["node_count"] == 0
    empty.close()


@pytest.mark.parametrize("kind", [sq_amd.AABB, sq_amd.SPHERE])
def test_f64_ties_coincident_parallel_and_unbounded(ctx, kind):
    """Over-full leaves from coincident centroids (tree.rs:145-148), exact
    duplicates (toi ties broken by leaf id), axis-parallel rays (the slab's
    parallel branch) and max_toi = +inf."""
    print(f"case: f64 ties and parallel rays kind={kind}", flush=True)
    rng = np.random.default_rng(0x7E5 + kind)
    dim = 3
    c = rng.uniform(-20, 20, (2000, dim))
    c[:300] = c[0]                      # 300 coincident centroids
    c[300:400] = c[300:400].round(0)    # duplicates on a lattice
    if kind == sq_amd.AABB:
        h = rng.uniform(0.2, 1.5, (2000, dim))
        h[:300] = rng.uniform(0.2, 1.5, (300, dim))
        prims = np.concatenate([c - h, c + h], 1)
        prims = np.concatenate([prims, prims[300:400]])           # exact copies
    else:
        r = rng.uniform(0.2, 1.5, (2000, 1))
        prims = np.concatenate([c, r], 1)
        prims = np.concatenate([prims, prims[300:400]])
    o = rng.uniform(-25, 25, (3000, dim))
    o[:200] = c[:200] + rng.uniform(-0.1, 0.1, (200, dim))       # start inside the coincident cluster
    d = rng.normal(size=(3000, dim))
    axis = rng.integers(0, dim, 1500)
    d[:1500] = 0.0
    d[np.arange(1500), axis] = rng.choice([-1.0, 1.0], 1500)    # axis-parallel rays
    d[1500:] /= np.linalg.norm(d[1500:], axis=1, keepdims=True)
    rays = np.concatenate([o, d], 1)
    b = sq_amd.DeviceBvh64(ctx, dim, kind, prims)
    ref = oracle.Scene64(dim, kind, prims)
    st, rs = b.stats(), ref.tree_stats()
    for k in ("node_count", "leaf_count", "prim_count", "max_depth", "bytes"):
        assert st[k] == rs[k], k
    for max_toi in (60.0, float("inf")):
        got = b.raycast_nearest(dev(rays), max_toi, normal=True)
        want = ref.raycast_nearest(rays, max_toi)
        assert (want["leaf"] >= 0).sum() > 1000
        assert np.array_equal(host(got["leaf"]).astype(np.int64), want["leaf"])
        assert np.array_equal(host(got["toi"]).view(np.uint64), want["t"].view(np.uint64))
        assert np.array_equal(host(got["normal"]).view(np.uint64), want["normal"].view(np.uint64))
        assert_csr(b.raycast_all(dev(rays[:1000]), max_toi), ref.raycast_all(rays[:1000], max_toi), "all, ties")
    casts = np.concatenate([rays, np.full((3000, dim), -0.25), np.full((3000, dim), 0.25)], 1)
    cn = b.shapecast_nearest(dev(casts), float("inf"))
    wc = ref.shapecast_nearest(casts, float("inf"))
    assert np.array_equal(host(cn["leaf"]).astype(np.int64), wc["leaf"])
    assert np.array_equal(host(cn["toi"]).view(np.uint64), wc["t"].view(np.uint64))
    regions = np.concatenate([c[:500] - 0.5, c[:500] + 0.5], 1)
    r_ov = ref.overlap(regions)
    assert_csr(b.overlap(dev(regions)), (r_ov[0], r_ov[1], None), "overlap, ties", has_t=False)
    b.close()


def orthonormal(rng, n, dim):
    q, r = np.linalg.qr(rng.normal(size=(n, dim, dim)))
    return q * np.sign(np.diagonal(r, axis1=1, axis2=2))[:, None, :]


@pytest.mark.parametrize("dim", [2, 3, 6])
def test_f64_box_instances_match_oracle(ctx, dim):
    """The f64 Tlas over box instances (two_level.rs:70-167 in double): nearest
    instance, toi and point bit for bit against the oracle's instance scan,
    any-hit, per-ray limits; the top tree is the f64 tree of the oracle's world
    boxes (same stats, so the world boxes agree)."""
    print(f"case: f64 box instances dim={dim}", flush=True)
    rng = np.random.default_rng(0xB0C5 + dim)
    n = 3000
    span = {2: 200.0, 3: 100.0, 6: 12.0}[dim]
    rot = orthonormal(rng, n, dim)
    trans = rng.uniform(-span, span, (n, dim))
    halves = rng.uniform(0.5, 4.0, (n, dim))
    o = rng.uniform(-span, span, (5000, dim))
    d = rng.normal(size=(5000, dim))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], 1)
    max_toi = 4.0 * span
    tl = sq_amd.Tlas64(ctx, dim, rot, trans, halves)
    ref = oracle.BoxInstances64(dim, rot, trans, halves)
    want = ref.raycast_nearest(rays, max_toi, world=True)
    assert (want["leaf"] >= 0).sum() > 800
    got = tl.raycast_nearest(dev(rays), max_toi, point=True)
    assert np.array_equal(host(got["leaf"]).astype(np.int64), want["leaf"])
    assert np.array_equal(host(got["toi"]).view(np.uint64), want["t"].view(np.uint64))
    hit = want["leaf"] >= 0
    pts = rays[:, :dim] + rays[:, dim:] * np.where(hit, want["t"], 0.0)[:, None]
    assert np.array_equal(host(got["point"])[hit].view(np.uint64), pts[hit].view(np.uint64))
    assert np.array_equal(host(tl.raycast_any(dev(rays), max_toi)).astype(bool), hit)
    lim = rng.uniform(0.1, max_toi, len(rays))
    got = tl.raycast_nearest(dev(rays), max_toi, max_toi_per_ray=dev(lim))
    per = [ref.raycast_nearest(rays[i:i + 1], lim[i]) for i in range(0, len(rays), 25)]
    idx = np.arange(0, len(rays), 25)
    assert np.array_equal(host(got["leaf"]).astype(np.int64)[idx], np.array([p["leaf"][0] for p in per]))
    assert np.array_equal(host(got["toi"])[idx].view(np.uint64), np.array([p["t"][0] for p in per]).view(np.uint64))
    st, rs = tl.stats(), oracle.Scene64(dim, oracle.BOX, want["world"]).tree_stats()
    for k in ("node_count", "leaf_count", "prim_count", "max_depth"):
        assert st[k] == rs[k], k
    assert st["quality"] == float(np.float32(rs["quality"]))
    tl.close()


@pytest.mark.parametrize("dim,scene", [(3, "identity"), (2, "identity"), (3, "plane")])
def test_f64_box_instances_on_their_world_faces(ctx, dim, scene):
    """f64 box instances whose faces lie on their world boxes (identity
    rotations) or that turn in one plane, non-dyadic centres and extents: the
    default mode equals the reference Tlas walk (unguarded instance test, the
    reference's order over the same f64 top tree) bit for bit; SQ_HITS_GUARDED
    equals the guarded oracle; the guard changes the answer on some rays."""
    print(f"case: f64 box instances D={dim}, {scene}, both hit modes", flush=True)
    rng = np.random.default_rng(0xFACE + dim)
    n, span = 2000, 60.0
    trans = rng.uniform(-span, span, (n, dim))
    halves = rng.uniform(0.3, 2.7, (n, dim))
    rot = np.broadcast_to(np.eye(dim), (n, dim, dim)).copy()
    if scene == "plane":
        ang = rng.uniform(0, 2 * np.pi, n)
        rot[:, 0, 0] = np.cos(ang); rot[:, 0, 1] = -np.sin(ang); rot[:, 1, 0] = np.sin(ang); rot[:, 1, 1] = np.cos(ang)
    o = rng.uniform(-span - 10, span + 10, (20000, dim))
    d = rng.normal(size=(20000, dim))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], 1)
    tl = sq_amd.Tlas64(ctx, dim, rot, trans, halves)
    plain = oracle.BoxInstances64(dim, rot, trans, halves).raycast_nearest(rays, 500.0)
    guarded = oracle.BoxInstances64(dim, rot, trans, halves, guarded=True).raycast_nearest(rays, 500.0)
    assert (plain["leaf"] >= 0).sum() > 2000
    got = tl.raycast_nearest(dev(rays), 500.0)
    assert np.array_equal(host(got["leaf"]).astype(np.int64), plain["leaf"])
    assert np.array_equal(host(got["toi"]).view(np.uint64), plain["t"].view(np.uint64))
    assert np.array_equal(host(tl.raycast_any(dev(rays), 500.0)).astype(bool), plain["leaf"] >= 0)
    try:
        ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED)
        g = tl.raycast_nearest(dev(rays), 500.0)
        assert np.array_equal(host(g["leaf"]).astype(np.int64), guarded["leaf"])
        assert np.array_equal(host(g["toi"]).view(np.uint64), guarded["t"].view(np.uint64))
    finally:
        ctx.set_hit_mode(sq_amd.SQ_HITS_REFERENCE)
    differ = int((plain["leaf"] != guarded["leaf"]).sum())
    print(f"rays whose reference answer the guard changes: {differ} of {len(rays)}", flush=True)
    assert differ > 0
    tl.close()


def test_f64_balls_both_hit_modes(ctx):
    """f64 balls in D = 2 (hits near box faces decide some rays): the default
    mode equals the reference's Bvh walk over the f64 tree, SQ_HITS_GUARDED the
    guarded oracle; nearest casts too."""
    print("case: f64 balls D=2, both hit modes", flush=True)
    rng = np.random.default_rng(0xBA2)
    prims = np.concatenate([rng.uniform(-40, 40, (8000, 2)), rng.uniform(0.3, 3.0, (8000, 1))], 1)
    o = rng.uniform(-45, 45, (40000, 2))
    d = rng.normal(size=(40000, 2))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], 1)
    b = sq_amd.DeviceBvh64(ctx, 2, sq_amd.SPHERE, prims)
    casts = np.concatenate([rays[:5000], np.full((5000, 2), -0.25), np.full((5000, 2), 0.25)], 1)
    for guarded in (False, True):
        ref = oracle.Scene64(2, sq_amd.SPHERE, prims, guarded=guarded)
        want = ref.raycast_nearest(rays, 300.0)
        wc = ref.shapecast_nearest(casts, 300.0)
        try:
            if guarded:
                ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED)
            got = b.raycast_nearest(dev(rays), 300.0)
            gc = b.shapecast_nearest(dev(casts), 300.0)
        finally:
            ctx.set_hit_mode(sq_amd.SQ_HITS_REFERENCE)
        assert np.array_equal(host(got["leaf"]).astype(np.int64), want["leaf"]), f"guarded={guarded}"
        assert np.array_equal(host(got["toi"]).view(np.uint64), want["t"].view(np.uint64))
        assert np.array_equal(host(gc["leaf"]).astype(np.int64), wc["leaf"]), f"casts guarded={guarded}"
        assert np.array_equal(host(gc["toi"]).view(np.uint64), wc["t"].view(np.uint64))
    b.close()


def _csr_equal64(got, want, with_t=True):
    off, leaf = host(got[0]).astype(np.uint64), host(got[1]).astype(np.int64)
    assert np.array_equal(off, want[0].astype(np.uint64)), "per-query counts differ"
    assert np.array_equal(leaf, want[1]), "ids or order differ"
    if with_t:
        assert np.array_equal(host(got[2]).view(np.uint64), want[2].view(np.uint64)), "toi bits differ"


@pytest.mark.parametrize("dim", [2, 3, 6])
def test_f64_tlas_box_instance_queries(ctx, dim):
    """The f64 Tlas over box instances answers every Tlas query as the f64
    oracle's Tlas walks (oracle.Instances64): all hits (= crossings) as CSR
    bit for bit in both hit modes, overlap (every instance of a leaf node
    touching the region, test_overlap's default), and shape casts that never
    touch a box instance (the trait default)."""
    print(f"case: f64 Tlas box-instance queries dim={dim}", flush=True)
    rng = np.random.default_rng(0x7A5 + dim)
    n = 3000
    span = {2: 200.0, 3: 100.0, 6: 12.0}[dim]
    rot = orthonormal(rng, n, dim)
    trans = rng.uniform(-span, span, (n, dim))
    halves = rng.uniform(0.5, 4.0, (n, dim))
    o = rng.uniform(-span, span, (4000, dim))
    d = rng.normal(size=(4000, dim))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], 1)
    max_toi = 4.0 * span
    tl = sq_amd.Tlas64(ctx, dim, rot, trans, halves)
    ref = oracle.Instances64(dim, 0, rot, trans, halves)
    want = ref.raycast_all(rays, max_toi)
    assert want[1].size > 1000
    for mode in (sq_amd.SQ_HITS_REFERENCE, sq_amd.SQ_HITS_GUARDED):
        ctx.set_hit_mode(mode)
        try:
            _csr_equal64(tl.raycast_all(dev(rays), max_toi), want)
        finally:
            ctx.set_hit_mode(sq_amd.SQ_HITS_REFERENCE)
    c = rng.uniform(-span, span, (3000, dim)); h = rng.uniform(0.05, 0.05 * span, (3000, dim))
    regions = np.concatenate([c - h, c + h], 1)
    wo = ref.overlap(regions)
    assert wo[1].size > 500
    _csr_equal64(tl.overlap(dev(regions)), wo, with_t=False)
    casts = np.concatenate([rays[:500], np.full((500, dim), -0.5), np.full((500, dim), 0.5)], 1)
    cn = tl.shapecast_nearest(dev(casts), max_toi, point=True)
    assert (host(cn["leaf"]) == -1).all() and np.isinf(host(cn["toi"])).all()
    off, leaf, _ = tl.shapecast_all(dev(casts), max_toi)
    assert int(host(off)[-1]) == 0
    tl.close()


@pytest.mark.parametrize("rotated", [False, True])
def test_f64_tlas_ball_instances(ctx, rotated):
    """Ball instances under the `f64` feature (shape_overlap.rs:288-345 in
    double): nearest rays, any-hit, all hits, shape casts (nearest and all)
    and exact overlap equal the oracle's Tlas walks bit for bit in the
    default mode, and the guarded oracle in SQ_HITS_GUARDED; with translations
    only they equal the naive scans over the same balls as a single-level
    provider (the reference's own tests' ground truth)."""
    print(f"case: f64 ball instances rotated={rotated}", flush=True)
    rng = np.random.default_rng(0xBA1 + rotated)
    n = 4000
    trans = rng.uniform(-60.0, 60.0, (n, 3))
    radii = rng.uniform(0.5, 2.0, n)
    rot = orthonormal(rng, n, 3) if rotated else np.tile(np.eye(3), (n, 1, 1))
    o = rng.uniform(-70.0, 70.0, (6000, 3))
    aim = trans[rng.integers(0, n, 6000)] + rng.uniform(-2.5, 2.5, (6000, 3)) - o
    rays = np.concatenate([o, aim / np.linalg.norm(aim, axis=1, keepdims=True)], 1)
    casts = np.concatenate([rays, np.full((6000, 3), -0.75), np.full((6000, 3), 0.75)], 1)
    c = rng.uniform(-60.0, 60.0, (3000, 3)); h = rng.uniform(1.0, 6.0, (3000, 3))
    regions = np.concatenate([c - h, c + h], 1)
    tl = sq_amd.Tlas64(ctx, 3, rot, trans, radii, kind=2)
    for guarded in (False, True):
        ref = oracle.Instances64(3, 1, rot, trans, radii, guarded=guarded)
        ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED if guarded else sq_amd.SQ_HITS_REFERENCE)
        try:
            w = ref.raycast_nearest(rays, 300.0)
            assert (w["leaf"] >= 0).sum() > 2000
            g = tl.raycast_nearest(dev(rays), 300.0)
            assert np.array_equal(host(g["leaf"]).astype(np.int64), w["leaf"])
            assert np.array_equal(host(g["toi"]).view(np.uint64), w["t"].view(np.uint64))
            assert np.array_equal(host(tl.raycast_any(dev(rays), 300.0)).astype(bool), w["leaf"] >= 0)
            wc = ref.shapecast_nearest(casts, 300.0)
            gc = tl.shapecast_nearest(dev(casts), 300.0)
            assert np.array_equal(host(gc["leaf"]).astype(np.int64), wc["leaf"])
            assert np.array_equal(host(gc["toi"]).view(np.uint64), wc["t"].view(np.uint64))
        finally:
            ctx.set_hit_mode(sq_amd.SQ_HITS_REFERENCE)
    ref = oracle.Instances64(3, 1, rot, trans, radii)
    _csr_equal64(tl.raycast_all(dev(rays), 300.0), ref.raycast_all(rays, 300.0))
    wa = ref.shapecast_all(casts, 300.0)
    assert wa[1].size > 2000
    _csr_equal64(tl.shapecast_all(dev(casts), 300.0), wa)
    wo = ref.overlap(regions)
    assert wo[1].size > 1000
    _csr_equal64(tl.overlap(dev(regions)), wo, with_t=False)
    if not rotated:
        balls = oracle.Scene64(3, oracle.SPHERE, np.concatenate([trans, radii[:, None]], 1), naive=True)
        wb = balls.raycast_nearest(rays, 300.0)
        g = tl.raycast_nearest(dev(rays), 300.0)
        assert np.array_equal(host(g["leaf"]).astype(np.int64), wb["leaf"])
        assert np.array_equal(host(g["toi"]).view(np.uint64), wb["t"].view(np.uint64))
        _csr_equal64(tl.overlap(dev(regions)), balls.overlap(regions), with_t=False)
        _csr_equal64(tl.shapecast_all(dev(casts), 300.0), balls.shapecast_all(casts, 300.0))
    tl.close()


# --- consumer narrow phases in double (sq_narrow64) --------------------------

def _consumer_scene(dim, n, seed, shrink=1.0):
    rng = np.random.default_rng(seed)
    span = {2: 60.0, 3: 30.0, 5: 9.0}[dim]
    c = rng.uniform(-span, span, (n, dim))
    r = rng.uniform(0.3, 3.0, (n, 1))
    o = rng.uniform(-span - 5.0, span + 5.0, (2 * n, dim))
    d = rng.normal(size=o.shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([c, r], 1), np.concatenate([o, d], 1), np.concatenate([c - shrink * r, c + shrink * r], 1)


def _same_nearest64(got, leaf, t, what):
    assert np.array_equal(host(got["leaf"]).astype(np.int64), np.asarray(leaf).astype(np.int64)), f"{what}: ids"
    assert np.array_equal(host(got["toi"]).view(np.uint64), np.asarray(t).view(np.uint64)), f"{what}: toi bits"


@pytest.mark.parametrize("dim", [3, 5])
def test_f64_consumer_ball_equals_builtin_sphere(ctx, tmp_path, dim):
    """The built-in ball's arithmetic as an f64 consumer (hipRTC into the
    pooled f64 walk) equals the engine's f64 SPHERE kind and the oracle's f64
    consumer kind (same source built by g++) bit for bit: nearest hits in both
    hit modes, all hits."""
    from consumer_narrow import BALL64, compile_cpu64
    print(f"case: f64 consumer ball D={dim}", flush=True)
    balls, rays, boxes = _consumer_scene(dim, 20000, 0xE0 + dim)
    nw = sq_amd.Narrow64(ctx, dim, BALL64, dim + 1)
    cb = nw.build(boxes, balls)
    builtin = sq_amd.DeviceBvh64(ctx, dim, sq_amd.SPHERE, balls)
    _, addr = compile_cpu64(BALL64, dim, dim + 1, str(tmp_path))
    saved = ctx.hit_mode()
    try:
        for mode, guarded in ((sq_amd.SQ_HITS_REFERENCE, False), (sq_amd.SQ_HITS_GUARDED, True)):
            ctx.set_hit_mode(mode)
            got = cb.raycast_nearest(dev(rays), 200.0)
            want = builtin.raycast_nearest(dev(rays), 200.0)
            _same_nearest64(got, host(want["leaf"]), host(want["toi"]), f"builtin, guarded={guarded}")
            ref = oracle.Consumer64(dim, boxes, balls, addr, guarded=guarded).raycast_nearest(rays, 200.0)
            assert (ref["leaf"] >= 0).sum() > 5000
            _same_nearest64(got, ref["leaf"], ref["t"], f"oracle, guarded={guarded}")
            ga = cb.raycast_all(dev(rays), 200.0)
            _csr_equal64(ga, tuple(host(x) for x in builtin.raycast_all(dev(rays), 200.0)))
            _csr_equal64(ga, oracle.Consumer64(dim, boxes, balls, addr).raycast_all(rays, 200.0))
        assert cb.stats() == builtin.stats()
    finally:
        ctx.set_hit_mode(saved)
        cb.close()
        builtin.close()
        nw.close()


@pytest.mark.parametrize("source", ["ball", "ellipsoid"])
def test_f64_consumer_hits_outside_their_boxes(ctx, tmp_path, source):
    """Providers whose hits leave their own world boxes (balls and ellipsoids
    whose boxes are shrunk to 0.6 of their extent): the default mode returns
    the reference's LIFO walk's answer (oracle f64 consumer kind, same source
    by g++) bit for bit, which differs from the naive scan on many rays;
    SQ_HITS_GUARDED returns the guarded oracle's; all hits equal the oracle's
    walk at max_toi."""
    from consumer_narrow import BALL64, ELLIPSOID64, compile_cpu64
    print(f"case: f64 consumer {source} with shrunk boxes", flush=True)
    dim = 3
    balls, rays, boxes = _consumer_scene(dim, 20000, 0xE7, shrink=0.6)
    if source == "ball":
        src, prims, pd = BALL64, balls, 4
    else:
        rng = np.random.default_rng(0xE8)
        semi = rng.uniform(0.3, 3.0, (balls.shape[0], 3))
        c = balls[:, :3]
        src, prims, pd = ELLIPSOID64, np.concatenate([c, semi], 1), 6
        boxes = np.concatenate([c - 0.6 * semi, c + 0.6 * semi], 1)
    nw = sq_amd.Narrow64(ctx, dim, src, pd)
    cb = nw.build(boxes, prims)
    _, addr = compile_cpu64(src, dim, pd, str(tmp_path))
    saved = ctx.hit_mode()
    try:
        ref = oracle.Consumer64(dim, boxes, prims, addr).raycast_nearest(rays, 200.0)
        naive = oracle.Consumer64(dim, boxes, prims, addr, naive=True).raycast_nearest(rays, 200.0)
        assert (ref["leaf"] != naive["leaf"]).sum() > 50
        ctx.set_hit_mode(sq_amd.SQ_HITS_REFERENCE)
        _same_nearest64(cb.raycast_nearest(dev(rays), 200.0), ref["leaf"], ref["t"], "reference order")
        ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED)
        g = oracle.Consumer64(dim, boxes, prims, addr, guarded=True).raycast_nearest(rays, 200.0)
        _same_nearest64(cb.raycast_nearest(dev(rays), 200.0), g["leaf"], g["t"], "guarded")
        _csr_equal64(cb.raycast_all(dev(rays), 200.0), oracle.Consumer64(dim, boxes, prims, addr).raycast_all(rays, 200.0))
    finally:
        ctx.set_hit_mode(saved)
        cb.close()
        nw.close()


def test_f64_consumer_errors_and_empty_tree(ctx):
    from consumer_narrow import BALL64, ELLIPSOID64
    print("case: f64 consumer compile error, refusals, empty tree", flush=True)
    with pytest.raises(sq_amd.SqError) as e:
        sq_amd.Narrow64(ctx, 3, "__device__ bool sq_narrow64(const double* p) { return undefined_name; }", 4)
    assert "error" in str(e.value)
    nw = sq_amd.Narrow64(ctx, 3, BALL64, 4)
    empty = nw.build(np.zeros((0, 6)), np.zeros((0, 4)))
    rays = _consumer_scene(3, 100, 5)[1]
    r = empty.raycast_nearest(dev(rays), 100.0)
    assert (host(r["leaf"]) == -1).all() and np.isinf(host(r["toi"])).all()
    off, leaf, _ = empty.raycast_all(dev(rays), 100.0)
    assert (host(off) == 0).all() and leaf.numel() == 0
    balls, rays, boxes = _consumer_scene(3, 50, 6)
    with pytest.raises(sq_amd.SqError):
        nw.build(np.concatenate([boxes[:, 3:], boxes[:, :3]], 1), balls)   # min > max
    cb = nw.build(boxes, balls)
    # a consumer tree answers only through its narrow phase
    bogus = sq_amd.DeviceBvh64.__new__(sq_amd.DeviceBvh64)
    bogus._h, bogus.ctx, bogus.dim, bogus.kind, bogus.n_prims = cb._h, ctx, 3, sq_amd.SPHERE, 50
    with pytest.raises(sq_amd.SqError):
        bogus.raycast_nearest(dev(rays), 100.0)
    bogus._h = None
    other = sq_amd.Narrow64(ctx, 3, ELLIPSOID64, 6)
    wrong = sq_amd.ConsumerBvh64(other, cb._h, 50)        # the tree under another narrow phase
    with pytest.raises(sq_amd.SqError):
        wrong.raycast_nearest(dev(rays), 100.0)
    wrong._h = None
    cb.close()
    empty.close()
    other.close()
    nw.close()


# --- refit under the feature (Bvh::refit, Tlas::refit, Tlas::needs_rebuild) ---

@pytest.mark.parametrize("dim,kind", [(3, oracle.SPHERE), (2, oracle.BOX)])
def test_f64_bvh_refit(ctx, dim, kind):
    """sq_bvh64_refit (bvh.rs:45-48, tree.rs:54-75 in double): built on one
    set of primitives and refit to moved ones, the tree answers nearest hits
    bit for bit as the oracle's tree built on the first set and refit to the
    second, in both hit modes, all hits as the naive scan over the moved set,
    and reports the refit tree's surface sum; a second refit back to the
    build's primitives gives the build's answers again."""
    print(f"case: f64 bvh refit dim={dim} kind={kind}", flush=True)
    rng = np.random.default_rng(0x5EF1 + dim)
    n = 20000
    span = {2: 60.0, 3: 30.0}[dim]
    c = rng.uniform(-span, span, (n, dim))
    p0 = np.concatenate([c, c + rng.uniform(0.2, 2.0, (n, dim))], 1) if kind == oracle.BOX else \
        np.concatenate([c, rng.uniform(0.2, 1.5, (n, 1))], 1)
    p2 = p0.copy()
    mv = rng.uniform(-3.0, 3.0, (n, dim))
    p2[:, :dim] += mv
    if kind == oracle.BOX:
        p2[:, dim:] += mv
    o = rng.uniform(-span, span, (20000, dim)); d = rng.normal(size=(20000, dim))
    rays = np.concatenate([o, d / np.linalg.norm(d, axis=1, keepdims=True)], 1)
    b = sq_amd.DeviceBvh64(ctx, dim, kind, p0)
    built = b.raycast_nearest(dev(rays), 200.0)
    b.refit(p2)
    for guarded in (False, True):
        ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED if guarded else sq_amd.SQ_HITS_REFERENCE)
        try:
            want = oracle.Scene64(dim, kind, p2, guarded=guarded).bvh_raycast_nearest_refit(p0, rays, 200.0, threads=8)
            if guarded:   # the guarded walk is order-independent: the naive guarded scan
                want = oracle.Scene64(dim, kind, p2, guarded=True, naive=True).raycast_nearest(rays, 200.0)
            got = b.raycast_nearest(dev(rays), 200.0)
            assert (want["leaf"] >= 0).sum() > 3000
            assert np.array_equal(host(got["leaf"]).astype(np.int64), want["leaf"])
            assert np.array_equal(host(got["toi"]).view(np.uint64), want["t"].view(np.uint64))
        finally:
            ctx.set_hit_mode(sq_amd.SQ_HITS_REFERENCE)
    q = oracle.Scene64(dim, kind, p2).bvh_raycast_nearest_refit(p0, rays[:1], 200.0)["quality"]
    assert b.stats()["quality"] == float(np.float32(q))
    _csr_equal64(b.raycast_all(dev(rays), 200.0), oracle.Scene64(dim, kind, p2).raycast_all(rays, 200.0))
    b.refit(p0)
    back = b.raycast_nearest(dev(rays), 200.0)
    assert np.array_equal(host(back["leaf"]), host(built["leaf"]))
    assert np.array_equal(host(back["toi"]).view(np.uint64), host(built["toi"]).view(np.uint64))
    with pytest.raises(ValueError):
        b.refit(p0[:-1])
    b.close()


@pytest.mark.parametrize("kind", [0, 2])
def test_f64_tlas_refit_and_needs_rebuild(ctx, kind):
    """sq_tlas64_refit / sq_tlas64_needs_rebuild (tlas.rs:49-60 in double): a
    Tlas over box (kind 0) or ball (kind 2) instances, refit to new rotations
    and translations, answers nearest hits in both hit modes and all hits bit
    for bit as the oracle's Tlas built on the first transforms and refit to
    the second; its surface sum is the oracle's, and needs_rebuild agrees with
    the oracle's at factors below and above the drift."""
    print(f"case: f64 tlas refit kind={kind}", flush=True)
    rng = np.random.default_rng(0x7EF1 + kind)
    n, dim = 3000, 3
    rot0 = orthonormal(rng, n, dim)
    t0 = rng.uniform(-60.0, 60.0, (n, dim))
    half = rng.uniform(0.5, 3.0, (n, dim)) if kind == 0 else rng.uniform(0.5, 2.0, n)
    rot1 = orthonormal(rng, n, dim)
    t1 = t0 + rng.uniform(-4.0, 4.0, (n, dim))
    o = rng.uniform(-70.0, 70.0, (8000, dim)); d = rng.normal(size=(8000, dim))
    rays = np.concatenate([o, d / np.linalg.norm(d, axis=1, keepdims=True)], 1)
    ok = 0 if kind == 0 else 1
    tl = sq_amd.Tlas64(ctx, dim, rot0, t0, half, kind=kind)
    q0 = oracle.Instances64(dim, ok, rot0, t0, half).quality()
    assert not tl.needs_rebuild(1.0)
    tl.refit(rot1, t1)   # halves kept
    for guarded in (False, True):
        ref = oracle.Instances64(dim, ok, rot1, t1, half, guarded=guarded, built_from=(rot0, t0, half))
        ctx.set_hit_mode(sq_amd.SQ_HITS_GUARDED if guarded else sq_amd.SQ_HITS_REFERENCE)
        try:
            w = ref.raycast_nearest(rays, 400.0)
            g = tl.raycast_nearest(dev(rays), 400.0)
            assert (w["leaf"] >= 0).sum() > 1000
            assert np.array_equal(host(g["leaf"]).astype(np.int64), w["leaf"])
            assert np.array_equal(host(g["toi"]).view(np.uint64), w["t"].view(np.uint64))
        finally:
            ctx.set_hit_mode(sq_amd.SQ_HITS_REFERENCE)
    ref = oracle.Instances64(dim, ok, rot1, t1, half, built_from=(rot0, t0, half))
    _csr_equal64(tl.raycast_all(dev(rays), 400.0), ref.raycast_all(rays, 400.0))
    q1 = ref.quality()
    assert tl.stats()["quality"] == float(np.float32(q1))
    for f in (1.0, q1 / q0 * 0.999, q1 / q0 * 1.001, 2.0):
        assert tl.needs_rebuild(f) == ref.needs_rebuild(f, q0), f
    tl.refit(rot0, t0, half)   # back to the build
    assert not tl.needs_rebuild(1.0)
    with pytest.raises(ValueError):
        tl.refit(rot0[:-1], t0[:-1])
    tl.close()


def test_f64_consumer_shell_crossings(ctx, tmp_path):
    """sq_raycast64_crossings_consumer (Bvh::raycast_crossings in double): a
    consumer whose sq_narrow64_crossings reports both roots of a shell gives
    the oracle's crossing lists (the same source built by g++) bit for bit, and
    its all-hits lists stay one hit per shell; a source without crossings
    answers crossings with its sq_narrow64 hits (the trait's default)."""
    from consumer_narrow import BALL64, SHELL64, compile_cpu64, crossings64_addr
    print("case: f64 consumer shell crossings", flush=True)
    rng = np.random.default_rng(0x5E12)
    n = 20000
    ctr = rng.uniform(-80.0, 80.0, (n, 3)); rad = rng.uniform(0.5, 3.0, (n, 1))
    prims = np.concatenate([ctr, rad], 1)
    boxes = np.concatenate([ctr - rad, ctr + rad], 1)
    o = rng.uniform(-90.0, 90.0, (20000, 3)); d = rng.normal(size=(20000, 3))
    rays = np.concatenate([o, d / np.linalg.norm(d, axis=1, keepdims=True)], 1)
    lib, addr = compile_cpu64(SHELL64, 3, 4, str(tmp_path))
    ref = oracle.Consumer64(3, boxes, prims, addr)
    nw = sq_amd.Narrow64(ctx, 3, SHELL64, 4)
    cb = nw.build(boxes, prims)
    want = ref.raycast_crossings(rays, 250.0, crossings64_addr(lib))
    assert want[1].size > 20000
    _csr_equal64(cb.raycast_crossings(dev(rays), 250.0), want)
    _csr_equal64(cb.raycast_all(dev(rays), 250.0), ref.raycast_all(rays, 250.0))
    cb.close(); nw.close()
    nb = sq_amd.Narrow64(ctx, 3, BALL64, 4)
    bb = nb.build(boxes, prims)
    _, baddr = compile_cpu64(BALL64, 3, 4, str(tmp_path))
    _csr_equal64(bb.raycast_crossings(dev(rays), 250.0), oracle.Consumer64(3, boxes, prims, baddr).raycast_all(rays, 250.0))
    bb.close(); nb.close()
