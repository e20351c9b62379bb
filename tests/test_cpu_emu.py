"""The engine's per-query logic (spatial-query_amd/csrc/traverse.h) compiled
for the CPU and run on the engine's own packed tree, checked against the
oracle.  Catches traversal and packing mistakes before any GPU run; the GPU
tests check the same logic on the device."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
from sq_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "cpu_emu", "_build", "libsqemu.so")
P = C.c_void_p


@pytest.fixture(scope="module")
def emu():
    if not os.path.exists(LIB):
        import subprocess
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpu_emu")], check=True)
    L = C.CDLL(LIB)
    L.emu_build.restype = P; L.emu_build.argtypes = [C.c_int, C.c_int, P, C.c_uint32]
    L.emu_free.argtypes = [P]
    L.emu_nearest.argtypes = [P, C.c_int, C.c_int, P, C.c_uint64, C.c_float, C.c_int, C.c_int, C.c_int, P, P, C.c_uint32,
                              C.c_int]
    L.emu_collect.restype = C.c_uint64
    L.emu_collect.argtypes = [P, C.c_int, C.c_int, P, C.c_uint64, C.c_float, C.c_int, P, P]
    L.emu_required.argtypes = [P, C.c_int, P, C.c_uint64, P, P, P]
    L.emu_nearest_quad.argtypes = [P, C.c_int, C.c_int, P, C.c_uint64, C.c_float, C.c_int, C.c_int, P, P, C.c_uint32,
                                   C.c_int, P]
    L.emu_nearest_pair_counted.argtypes = [P, C.c_int, C.c_int, P, C.c_uint64, C.c_float, C.c_int, C.c_uint32, P, P, P]
    L.emu_nearest_spec.argtypes = [P, C.c_int, C.c_int, P, C.c_uint64, C.c_float, C.c_int, C.c_uint32, C.c_uint32, P, P, P]
    L.emu_n_synth.restype = C.c_uint32
    L.emu_n_synth.argtypes = [P]
    L.emu_counted.argtypes = [P, C.c_int, C.c_int, P, C.c_uint64, C.c_float, C.c_int, P, P, P]
    return L


def p(a):
    return a.ctypes.data_as(P)


class Emu:
    def __init__(self, L, dim, kind, prims):
        self.L, self.dim, self.kind = L, dim, kind
        self.prims = np.ascontiguousarray(prims, np.float32)
        self.h = L.emu_build(dim, kind, p(self.prims), self.prims.shape[0])

    def nearest(self, q, max_toi, any_=0, cast=0, ring=0, guarded=False):
        q = np.ascontiguousarray(q, np.float32)
        leaf = np.empty(len(q), np.int32); t = np.empty(len(q), np.float32)
        self.L.emu_nearest(self.h, self.dim, self.kind, p(q), len(q), max_toi, any_, cast, 8, p(leaf), p(t), ring,
                           int(guarded))
        return leaf.astype(np.int64), t

    def nearest_quad(self, q, max_toi, any_=0, ring=8, guarded=True, counted=False):
        q = np.ascontiguousarray(q, np.float32)
        leaf = np.empty(len(q), np.int32); t = np.empty(len(q), np.float32)
        tested = np.zeros(len(q), np.uint32)
        rc = self.L.emu_nearest_quad(self.h, self.dim, self.kind, p(q), len(q), max_toi, any_, 8, p(leaf), p(t), ring,
                                     int(guarded), p(tested))
        assert rc == 0, rc
        return (leaf.astype(np.int64), t, tested) if counted else (leaf.astype(np.int64), t)

    def nearest_pair_counted(self, q, max_toi, ring=8, guarded=False):
        q = np.ascontiguousarray(q, np.float32)
        leaf = np.empty(len(q), np.int32); t = np.empty(len(q), np.float32)
        tested = np.zeros(len(q), np.uint32)
        rc = self.L.emu_nearest_pair_counted(self.h, self.dim, self.kind, p(q), len(q), max_toi, int(guarded), ring,
                                             p(leaf), p(t), p(tested))
        assert rc == 0, rc
        return leaf.astype(np.int64), t, tested

    def nearest_spec(self, q, max_toi, quad, ring=8, pause_every=0):
        q = np.ascontiguousarray(q, np.float32)
        leaf = np.empty(len(q), np.int32); t = np.empty(len(q), np.float32)
        tested = np.zeros(len(q), np.uint32)
        rc = self.L.emu_nearest_spec(self.h, self.dim, self.kind, p(q), len(q), max_toi, int(quad), ring, pause_every,
                                     p(leaf), p(t), p(tested))
        assert rc == 0, rc
        return leaf.astype(np.int64), t, tested

    def counted(self, q, max_toi, all_=0):
        q = np.ascontiguousarray(q, np.float32)
        leaf = np.empty(len(q), np.int32); t = np.empty(len(q), np.float32)
        st = np.empty((len(q), 4), np.uint32)
        self.L.emu_counted(self.h, self.dim, self.kind, p(q), len(q), max_toi, all_, p(leaf), p(t), p(st))
        return leaf.astype(np.int64), t, st

    def collect(self, q, max_toi, mode):
        q = np.ascontiguousarray(q, np.float32)
        off = np.empty(len(q) + 1, np.uint64)
        tot = self.L.emu_collect(self.h, self.dim, self.kind, p(q), len(q), max_toi, mode, p(off), None)
        keys = np.empty(tot, np.uint64)
        self.L.emu_collect(self.h, self.dim, self.kind, p(q), len(q), max_toi, mode, p(off), p(keys))
        for i in range(len(q)):
            keys[off[i]:off[i + 1]] = np.sort(keys[off[i]:off[i + 1]])
        return off, (keys & np.uint64(0xFFFFFFFF)).astype(np.int64)

    def __del__(self):
        self.L.emu_free(self.h)


def eq(a_leaf, a_t, b_leaf, b_t):
    return np.array_equal(a_leaf, b_leaf) and np.array_equal(a_t.view(np.uint32), b_t.view(np.uint32))


def test_c0_nearest_any_required(emu):
    sc, rays = W.c0()
    o = oracle.Scene(3, oracle.BOX, sc.prims)
    ref = o.raycast_nearest(rays, 400.0, threads=8)
    e = Emu(emu, 3, 0, sc.prims)
    assert eq(*e.nearest(rays, 400.0), ref["leaf"], ref["t"])
    assert np.array_equal(e.nearest(rays, 400.0, any_=1)[0] >= 0, ref["leaf"] >= 0)
    tl = np.where(ref["leaf"] >= 0, ref["t"], np.float32(400.0)).astype(np.float32)
    on = np.empty(len(rays), np.uint32); op = np.empty(len(rays), np.uint32)
    emu.emu_required(e.h, 3, p(rays), len(rays), p(tl), p(on), p(op))
    rn, rp = o.required_work(rays, tl, threads=8)
    assert np.array_equal(on, rn) and np.array_equal(op, rp)


@pytest.mark.parametrize("dim", [2, 3, 4, 6])
def test_balls_nearest(emu, dim):
    s = W.Stream(dim)
    prims = np.concatenate([s.range(-50.0, 50.0, 3000 * dim).reshape(-1, dim), s.range(0.5, 4.0, 3000)[:, None]], 1)
    rays = W.random_rays(20000, 77 + dim, -60.0, 60.0, dim=dim)
    ref = oracle.Scene(dim, oracle.SPHERE, prims).raycast_nearest(rays, 500.0, threads=8)
    assert eq(*Emu(emu, dim, 1, prims).nearest(rays, 500.0), ref["leaf"], ref["t"])


def test_heightfield_and_coincident_leaves(emu):
    sc = W.heightfield_scene(grid=257, extent=128.0)
    rays = np.concatenate([W.camera_rays(sc, res=256, seed=0x11), W.incoherent_rays(sc, 65536)])
    ref = oracle.Scene(3, oracle.TRI, sc.prims).raycast_nearest(rays, 2000.0, threads=8)
    assert (ref["leaf"] >= 0).sum() > 30000
    assert eq(*Emu(emu, 3, 2, sc.prims).nearest(rays, 2000.0), ref["leaf"], ref["t"])
    # 100 coincident balls -> one reference leaf, split into synthetic chains
    prims = np.tile(np.array([[5, 0, 0, 1]], np.float32), (100, 1))
    prims[::7, 3] = 1.5
    ray = np.array([[0, 0, 0, 1, 0, 0], [0, 0.5, 0, 1, 0, 0]], np.float32)
    o = oracle.Scene(3, oracle.SPHERE, prims)
    ref = o.raycast_nearest(ray, 100.0)
    e = Emu(emu, 3, 1, prims)
    assert eq(*e.nearest(ray, 100.0), ref["leaf"], ref["t"])
    off, leaf = e.collect(ray, 100.0, 0)
    roff, rleaf, _ = o.raycast_all(ray, 100.0)
    assert np.array_equal(off, roff)
    for i in range(2):   # same (toi, leaf) sets; the emu sorts keys
        assert np.array_equal(np.sort(leaf[off[i]:off[i + 1]]), np.sort(rleaf[roff[i]:roff[i + 1]]))


@pytest.mark.parametrize("dim", [3, 6])
def test_collect_and_casts(emu, dim):
    sc = W.boxes_scene(n=4000, dim=dim, lo=-30.0, hi=30.0, hmin=0.5, hmax=3.0)
    q = W.random_rays(2000, 5, -40.0, 40.0, dim=dim)
    o = oracle.Scene(dim, oracle.BOX, sc.prims)
    e = Emu(emu, dim, 0, sc.prims)
    for mode, cross in ((0, False), (1, True)):
        off, leaf = e.collect(q, 500.0, mode)
        roff, rleaf, rt = o.raycast_all(q, 500.0, crossings=cross)
        assert np.array_equal(off, roff)
        # keys sort by (toi, leaf): the leaf order then matches hit_ordering
        assert np.array_equal(leaf, rleaf)
    regions = np.concatenate([q[:, :dim] - 3, q[:, :dim] + 3], 1).astype(np.float32)
    off, leaf = e.collect(regions, 0.0, 2)
    roff, rleaf = o.overlap(regions)
    assert np.array_equal(off, roff) and np.array_equal(leaf, rleaf)
    half = np.float32(0.5)
    casts = np.concatenate([q, np.full((len(q), dim), -half), np.full((len(q), dim), half)], 1).astype(np.float32)
    ref = o.shapecast_nearest(casts, 50.0)
    assert eq(*e.nearest(casts, 50.0, cast=1), ref["leaf"], ref["t"])
    off, leaf = e.collect(casts, 50.0, 3)
    roff, rleaf, _ = o.shapecast_all(casts, 50.0)
    assert np.array_equal(off, roff) and np.array_equal(leaf, rleaf)


@pytest.mark.parametrize("ring", [2, 4, 8])
def test_ring_stack_walk(emu, ring):
    """The pooled kernel's resumable walk with an LDS ring of `ring` entries
    (2 spills on almost every push) equals the oracle."""
    sc = W.heightfield_scene(grid=257, extent=128.0)
    rays = np.concatenate([W.camera_rays(sc, res=128, seed=0x11), W.incoherent_rays(sc, 30000)])
    ref = oracle.Scene(3, oracle.TRI, sc.prims).raycast_nearest(rays, 2000.0, threads=8)
    e = Emu(emu, 3, 2, sc.prims)
    assert eq(*e.nearest(rays, 2000.0, ring=ring), ref["leaf"], ref["t"])
    assert np.array_equal(e.nearest(rays, 2000.0, any_=1, ring=ring)[0] >= 0, ref["leaf"] >= 0)
    sc6 = W.boxes_scene(n=3000, dim=6, lo=-7.0, hi=7.0, hmin=0.3, hmax=2.0)
    q = W.random_rays(3000, 9, -17.0, 17.0, dim=6)
    o6 = oracle.Scene(6, oracle.BOX, sc6.prims)
    e6 = Emu(emu, 6, 0, sc6.prims)
    r6 = o6.raycast_nearest(q, 500.0)
    assert eq(*e6.nearest(q, 500.0, ring=ring), r6["leaf"], r6["t"])
    casts = np.concatenate([q, np.full((len(q), 6), -0.5, np.float32), np.full((len(q), 6), 0.5, np.float32)], 1)
    rc = o6.shapecast_nearest(casts, 50.0)
    assert eq(*e6.nearest(casts, 50.0, cast=1, ring=ring), rc["leaf"], rc["t"])


@pytest.mark.parametrize("ring", [2, 8])
def test_quad_walk(emu, ring):
    """The 4-wide walk (tuning `wide`: each pair node with its children's pair
    nodes; near-first, so for balls and triangles it runs under
    SQ_HITS_GUARDED) equals the oracle: heightfield triangles with coherent and
    incoherent rays, coincident boxes (leaf chains), balls in D = 2 and 5 and
    boxes in D = 6, nearest and any-hit, with rings that spill."""
    sc = W.heightfield_scene(grid=257, extent=128.0)
    rays = np.concatenate([W.camera_rays(sc, res=128, seed=0x12), W.incoherent_rays(sc, 30000)])
    ref = oracle.Scene(3, oracle.TRI, sc.prims, guarded=True).raycast_nearest(rays, 2000.0, threads=8)
    e = Emu(emu, 3, 2, sc.prims)
    assert eq(*e.nearest_quad(rays, 2000.0, ring=ring), ref["leaf"], ref["t"])
    assert np.array_equal(e.nearest_quad(rays, 2000.0, any_=1, ring=ring)[0] >= 0, ref["leaf"] >= 0)
    co = np.concatenate([np.tile(np.array([[1.0, 2.0, 3.0, 1.5, 2.5, 3.5]], np.float32), (1000, 1)),
                         W.boxes_scene(n=500, seed=5).prims])
    q = W.random_rays(5000, 0x13, -120.0, 120.0)
    rc = oracle.Scene(3, oracle.BOX, co).raycast_nearest(q, 500.0)
    assert eq(*Emu(emu, 3, 0, co).nearest_quad(q, 500.0, ring=ring), rc["leaf"], rc["t"])
    for dim, half in ((2, 40.0), (5, 10.0)):
        s = W.Stream(0x14 + dim)
        b = np.concatenate([s.range(-half, half, 4000 * dim).reshape(-1, dim), s.range(0.3, 3.0, 4000).reshape(-1, 1)],
                           1).astype(np.float32)
        qq = W.random_rays(4000, 0x15 + dim, -half - 5.0, half + 5.0, dim=dim)
        rb = oracle.Scene(dim, oracle.SPHERE, b, guarded=True).raycast_nearest(qq, 300.0)
        assert (rb["leaf"] >= 0).sum() > 100
        assert eq(*Emu(emu, dim, 1, b).nearest_quad(qq, 300.0, ring=ring), rb["leaf"], rb["t"])
    sc6 = W.boxes_scene(n=3000, dim=6, lo=-7.0, hi=7.0, hmin=0.3, hmax=2.0)
    q6 = W.random_rays(3000, 9, -17.0, 17.0, dim=6)
    r6 = oracle.Scene(6, oracle.BOX, sc6.prims).raycast_nearest(q6, 500.0)
    assert eq(*Emu(emu, 6, 0, sc6.prims).nearest_quad(q6, 500.0, ring=ring), r6["leaf"], r6["t"])
    one = Emu(emu, 3, 0, W.boxes_scene(n=1, seed=1).prims)
    r1 = oracle.Scene(3, oracle.BOX, W.boxes_scene(n=1, seed=1).prims).raycast_nearest(q, 500.0)
    assert eq(*one.nearest_quad(q, 500.0, ring=ring), r1["leaf"], r1["t"])


@pytest.mark.parametrize("ring", [2, 8])
def test_quad_walk_reference_order(emu, ring):
    """The 4-wide walk in the reference's order (SQ_HITS_REFERENCE, plain
    narrow phases, trees without synthetic pairs) equals the oracle's plain
    providers on every ray, including the rays whose answer is a hit outside
    its own box (where any other order may differ): the heightfield with
    coherent and incoherent rays, balls in D = 2 and 5, nearest and any-hit.
    A walk order that takes child a's subtree first, or a grandchild tested
    at push time only, fails on those rays."""
    sc = W.heightfield_scene(grid=257, extent=128.0)
    rays = np.concatenate([W.camera_rays(sc, res=128, seed=0x12), W.incoherent_rays(sc, 30000)])
    plain = oracle.Scene(3, oracle.TRI, sc.prims).raycast_nearest(rays, 2000.0, threads=8)
    guarded = oracle.Scene(3, oracle.TRI, sc.prims, guarded=True).raycast_nearest(rays, 2000.0, threads=8)
    e = Emu(emu, 3, 2, sc.prims)
    assert emu.emu_n_synth(e.h) == 0
    ql, qt, qn = e.nearest_quad(rays, 2000.0, ring=ring, guarded=False, counted=True)
    assert eq(ql, qt, plain["leaf"], plain["t"])
    # the same leaves in the same order with the same limits: every ray tests
    # exactly the primitives the pair walk in the reference's order tests
    # (a near-first or a-first order tests different ones on most hitting rays)
    pl, pt, pn = e.nearest_pair_counted(rays, 2000.0, ring=ring, guarded=False)
    assert eq(pl, pt, plain["leaf"], plain["t"])
    assert np.array_equal(qn, pn)
    gn = e.nearest_quad(rays, 2000.0, ring=ring, guarded=True, counted=True)[2]
    assert (gn != pn).sum() > len(rays) // 10
    assert np.array_equal(e.nearest_quad(rays, 2000.0, any_=1, ring=ring, guarded=False)[0] >= 0, plain["leaf"] >= 0)
    differ = int((plain["leaf"] != guarded["leaf"]).sum())
    for dim, half in ((2, 40.0), (5, 10.0)):
        s = W.Stream(0x16 if dim == 2 else 0x14 + dim)
        b = np.concatenate([s.range(-half, half, 4000 * dim).reshape(-1, dim), s.range(0.3, 3.0, 4000).reshape(-1, 1)],
                           1).astype(np.float32)
        qq = W.random_rays(4000, 0x17 if dim == 2 else 0x15 + dim, -half - 5.0, half + 5.0, dim=dim)
        eb = Emu(emu, dim, 1, b)
        if emu.emu_n_synth(eb.h):
            continue
        rb = oracle.Scene(dim, oracle.SPHERE, b).raycast_nearest(qq, 300.0)
        rg = oracle.Scene(dim, oracle.SPHERE, b, guarded=True).raycast_nearest(qq, 300.0)
        differ += int((rb["leaf"] != rg["leaf"]).sum())
        ql, qt, qn = eb.nearest_quad(qq, 300.0, ring=ring, guarded=False, counted=True)
        assert eq(ql, qt, rb["leaf"], rb["t"])
        assert np.array_equal(qn, eb.nearest_pair_counted(qq, 300.0, ring=ring)[2])
        assert np.array_equal(eb.nearest_quad(qq, 300.0, any_=1, ring=ring, guarded=False)[0] >= 0, rb["leaf"] >= 0)
    assert differ > 0, "the scenes should hold rays whose reference answer lies outside its box"


@pytest.mark.parametrize("ring,pause_every", [(2, 0), (8, 0), (8, 3), (4, 7)])
def test_lookahead_walk(emu, ring, pause_every):
    """The look-ahead form of the reference-order walk (walk_step_spec: a
    lane keeps a reached leaf pending and descends on to its next leaf with
    the limit it has, then tests the pending leaf and checks its position
    again) tests exactly the primitives the pair walk in the reference's order
    tests, ray for ray, and returns the oracle's plain answers: over the 4-wide
    records and over the pair nodes, on the heightfield (camera and incoherent
    rays), balls in D = 2 and 5, and coincident balls (an over-full leaf walked
    as synthetic pairs), with descents paused and resumed and rings that
    spill."""
    sc = W.heightfield_scene(grid=257, extent=128.0)
    rays = np.concatenate([W.camera_rays(sc, res=128, seed=0x12), W.incoherent_rays(sc, 30000)])
    cases = [(3, 2, oracle.TRI, sc.prims, rays, 2000.0)]
    for dim, half in ((2, 40.0), (5, 10.0)):
        s = W.Stream(0x16 if dim == 2 else 0x14 + dim)
        b = np.concatenate([s.range(-half, half, 4000 * dim).reshape(-1, dim), s.range(0.3, 3.0, 4000).reshape(-1, 1)],
                           1).astype(np.float32)
        cases.append((dim, 1, oracle.SPHERE, b, W.random_rays(4000, 0x17 if dim == 2 else 0x15 + dim, -half - 5.0,
                                                              half + 5.0, dim=dim), 300.0))
    dup = np.concatenate([np.tile(np.array([[5, 0, 0, 1]], np.float32), (100, 1)),
                          np.tile(np.array([[-5, 3, 1, 2]], np.float32), (60, 1))])
    dup[::7, 3] = 1.5
    cases.append((3, 1, oracle.SPHERE, dup, W.random_rays(3000, 0x19, -8.0, 8.0), 100.0))
    synth_seen = False
    for dim, kind, ok, prims, q, tmax in cases:
        e = Emu(emu, dim, kind, prims)
        plain = oracle.Scene(dim, ok, prims).raycast_nearest(q, tmax, threads=8)
        pl, pt, pn = e.nearest_pair_counted(q, tmax, ring=ring)
        assert eq(pl, pt, plain["leaf"], plain["t"])
        quads = (True, False) if emu.emu_n_synth(e.h) == 0 else (False,)
        synth_seen |= emu.emu_n_synth(e.h) > 0
        for quad in quads:
            sl, stt, sn = e.nearest_spec(q, tmax, quad, ring=ring, pause_every=pause_every)
            assert eq(sl, stt, plain["leaf"], plain["t"]), (dim, kind, quad)
            assert np.array_equal(sn, pn), (dim, kind, quad, int((sn != pn).sum()))
    assert synth_seen


@pytest.mark.parametrize("ring", [0, 2, 8])
def test_hit_modes(emu, ring):
    """Both hit modes of the walk against the oracle on scenes where they
    differ: SQ_HITS_REFERENCE (plain narrow phases, the reference's order)
    equals the oracle's plain providers, and SQ_HITS_GUARDED (guarded narrow
    phases, near-first) the guarded ones, for balls in D = 2 (where a hit
    outside its box's slab interval decides some rays) and the heightfield;
    any-hit is the plain existence of a hit in the reference mode."""
    s = W.Stream(0x16)
    b = np.concatenate([s.range(-40.0, 40.0, 4000 * 2).reshape(-1, 2), s.range(0.3, 3.0, 4000).reshape(-1, 1)],
                       1).astype(np.float32)
    q = W.random_rays(4000, 0x17, -45.0, 45.0, dim=2)
    sc = W.heightfield_scene(grid=257, extent=128.0)
    rays = np.concatenate([W.camera_rays(sc, res=128, seed=0x12), W.incoherent_rays(sc, 30000)])
    differ = 0
    for dim, kind, ok, prims, qq, tmax in ((2, 1, oracle.SPHERE, b, q, 300.0), (3, 2, oracle.TRI, sc.prims, rays, 2000.0)):
        e = Emu(emu, dim, kind, prims)
        plain = oracle.Scene(dim, ok, prims).raycast_nearest(qq, tmax, threads=8)
        guarded = oracle.Scene(dim, ok, prims, guarded=True).raycast_nearest(qq, tmax, threads=8)
        assert eq(*e.nearest(qq, tmax, ring=ring), plain["leaf"], plain["t"])
        assert eq(*e.nearest(qq, tmax, ring=ring, guarded=True), guarded["leaf"], guarded["t"])
        assert np.array_equal(e.nearest(qq, tmax, any_=1, ring=ring)[0] >= 0, plain["leaf"] >= 0)
        differ += int((plain["leaf"] != guarded["leaf"]).sum())
    assert differ > 0, "the scenes should hold rays where the two modes differ"


def _counted_cases():
    """(name, dim, kind, prims, rays, max_toi): boxes, balls D = 2..6, the
    heightfield, and 100 coincident balls (one over-full reference leaf)."""
    out = []
    sc = W.boxes_scene(n=3000, seed=0xC0C0, lo=-100.0, hi=100.0, hmin=0.1, hmax=2.0)
    out.append(("boxes", 3, 0, sc.prims, W.random_rays(8000, 0xC0C1, -150.0, 150.0), 400.0))
    for dim in (2, 4, 6):
        s = W.Stream(40 + dim)
        prims = np.concatenate([s.range(-50.0, 50.0, 2000 * dim).reshape(-1, dim), s.range(0.5, 4.0, 2000)[:, None]], 1)
        out.append((f"balls{dim}", dim, 1, prims, W.random_rays(6000, 90 + dim, -60.0, 60.0, dim=dim), 500.0))
    hf = W.heightfield_scene(grid=129, extent=64.0)
    rays = np.concatenate([W.camera_rays(hf, res=64, seed=0x12), W.incoherent_rays(hf, 8192)])
    out.append(("heightfield", 3, 2, hf.prims, rays, 2000.0))
    dup = np.tile(np.array([[5, 0, 0, 1]], np.float32), (100, 1))
    dup[::7, 3] = 1.5
    out.append(("coincident", 3, 1, dup, np.array([[0, 0, 0, 1, 0, 0], [0, 0.5, 0, 1, 0, 0], [0, 9, 0, 1, 0, 0]],
                                                  np.float32), 100.0))
    return out


@pytest.mark.parametrize("case", range(6))
def test_reference_counters(emu, case):
    """The device's counted walk (walk_counted, run on the CPU) gives the
    reference's QueryStats on every ray, nearest and all hits, and the
    reference's nearest hit."""
    name, dim, kind, prims, rays, tmax = _counted_cases()[case]
    okind = {0: oracle.BOX, 1: oracle.SPHERE, 2: oracle.TRI}[kind]
    o = oracle.Scene(dim, okind, prims)
    e = Emu(emu, dim, kind, prims)
    ref = o.raycast_nearest(rays, tmax, threads=8, stats=True)
    leaf, t, st = e.counted(rays, tmax)
    assert eq(leaf, t, ref["leaf"], ref["t"]), name
    assert np.array_equal(st.astype(np.uint64), ref["stats"]), name
    _, _, sa = e.counted(rays, tmax, all_=1)
    assert np.array_equal(sa.astype(np.uint64), o.raycast_all_stats(rays, tmax)), name
    if name == "coincident":     # one reference leaf of 100 balls: one node, every ball tested
        assert ref["stats"][:, 0].tolist() == [1, 1, 1] and ref["stats"][:2, 2].tolist() == [100, 100]
    else:
        assert (ref["stats"][:, 0] > 1).any()
