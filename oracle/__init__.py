"""ctypes binding of the CPU oracle (sq_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by `tests/`, by the smoke check and by the
`cpu_baseline` leg of `bench.py`, always as the checker or the timed CPU
baseline, never as a code path of the product.

The oracle is a C++ restatement of the CPU path of grimandgreedy/spatial-query
(see the header of sq_oracle.cpp for the file:line map).  Build it with
`make -C oracle` (done by `__graft_entry__.build()`).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")

# Provider kinds (sq_oracle.cpp `enum Kind`).
BOX, SPHERE, TRI = 0, 1, 2
SPHERE_PLAIN, SPHERE_PLAIN_NOEXIT, TRI_PLAIN, SHELL = 10, 11, 12, 13
KIND_CONSUMER = 20
INST_BOX, INST_MESH, INST_BOX_GUARDED, INST_BALL = 0, 1, 2, 3
ALL_LAYERS = (1 << 64) - 1

_lib = None


def build(force: bool = False) -> None:
    """Compile the oracle with its Makefile (g++ -O3 -ffp-contract=off)."""
    if force and os.path.exists(_LIB_PATH):
        os.remove(_LIB_PATH)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


P = C.c_void_p
U64, I64, F32, I32, U32 = C.c_uint64, C.c_int64, C.c_float, C.c_int, C.c_uint32


def _declare(L):
    sig = {
        "orc_rng_range": (None, [U64, F32, F32, U64, P]),
        "orc_rng_u64": (None, [U64, U64, P]),
        "orc_normalize": (None, [I32, P, U64]),
        "orc_geom_new": (P, [I32, I32, P, U32]),
        "orc_geom_set_guarded": (None, [P, I32]),
        "orc_inst_set_guarded": (None, [P, I32]),
        "orc_geom_new_consumer": (P, [I32, P, U32, I32, U64]),
        "orc_geom_set_consumer_crossings": (None, [P, U64]),
        "orc_geom_set_consumer_cast_overlap": (None, [P, U64, U64]),
        "orc_geom_set_layers": (None, [P, P]),
        "orc_geom_set_prims": (None, [P, P]),
        "orc_geom_world_aabbs": (None, [P, P]),
        "orc_geom_free": (None, [P]),
        "orc_bvh_build": (P, [P]),
        "orc_bvh_refit": (None, [P, P]),
        "orc_bvh_node_count": (U64, [P]),
        "orc_bvh_export": (None, [P, P, P, P, P, P]),
        "orc_bvh_stats": (None, [P, P, P]),
        "orc_bvh_free": (None, [P]),
        "orc_raycast_nearest_batch": (None, [P, P, P, U64, F32, U64, I64, I32, I32, P, P, P, P, P, P]),
        "orc_raycast_all_batch": (U64, [P, P, P, U64, F32, U64, I64, I32, I32, P, P, P, U64]),
        "orc_raycast_all_stats": (None, [P, P, P, U64, F32, U64, I64, P]),
        "orc_overlap_batch": (U64, [P, P, P, U64, U64, I64, I32, P, P, U64]),
        "orc_shapecast_nearest_batch": (None, [P, P, P, U64, P, U64, I64, I32, P, P, P, P]),
        "orc_shapecast_all_batch": (U64, [P, P, P, U64, P, U64, I64, I32, P, P, P, U64]),
        "orc_required_work": (None, [P, P, U64, P, P, P, I32]),
        "orc64_raycast_nearest": (None, [I32, I32, P, U32, P, U64, I64, P, U64, C.c_double, P, I32, P, P, P, P]),
        "orc64_raycast_all": (U64, [I32, I32, P, U32, P, U64, I64, P, U64, C.c_double, I32, P, P, P, U64]),
        "orc64_overlap": (U64, [I32, I32, P, U32, P, U64, I64, P, U64, I32, P, P, U64]),
        "orc64_tree_stats": (None, [I32, I32, P, U32, P, P]),
        "orc64_consumer_query": (U64, [I32, I32, P, U32, I32, U64, P, U64, C.c_double, I32, I32, P, P, P, U64]),
        "orc64_consumer_crossings": (U64, [I32, I32, P, U32, I32, U64, U64, P, U64, C.c_double, I32, P, P, P, U64]),
        "orc64_bvh_raycast_nearest": (None, [I32, I32, P, U32, P, U64, C.c_double, I32, P, P, P]),
        "orc64_bvh_refit_raycast_nearest": (None, [I32, I32, P, P, U32, P, U64, C.c_double, I32, P, P, P, P]),
        "orc64_tlas_refit_query": (U64, [I32, I32, U64, P, P, P, P, P, P, P, U64, C.c_double, I32, I32, I32, P, P, P, U64,
                                          P]),
        "orc64_tlas_box_nearest": (None, [I32, U64, P, P, P, P, U64, C.c_double, I32, I32, P, P, P]),
        "orc64_tlas_query": (U64, [I32, I32, U64, P, P, P, P, U64, C.c_double, I32, I32, I32, P, P, P, U64]),
        "orc64_shapecast_nearest": (None, [I32, I32, P, U32, P, U64, I64, P, U64, C.c_double, P, I32, P, P, P]),
        "orc64_shapecast_all": (U64, [I32, I32, P, U32, P, U64, I64, P, U64, C.c_double, P, I32, P, P, P, U64]),
        "orc_inst_new": (P, [I32, I32, U64, P, P, P, U32, P, P, P, P]),
        "orc_inst_set_transforms": (None, [P, P, P, P]),
        "orc_tlas_refit": (None, [P]),
        "orc_tlas_rebuild": (None, [P]),
        "orc_tlas_needs_rebuild": (I32, [P, F32]),
        "orc_tlas_node_count": (U64, [P]),
        "orc_tlas_export": (None, [P, P, P, P, P, P, P]),
        "orc_tlas_raycast_nearest_batch": (None, [P, P, U64, F32, I32, I32, P, P, P, P, P]),
        "orc_tlas_raycast_all_batch": (U64, [P, P, U64, F32, P, P, P, P, U64]),
        "orc_tlas_overlap_batch": (U64, [P, P, U64, P, P, U64]),
        "orc_tlas_shapecast_batch": (U64, [P, P, U64, F32, I32, P, P, P, U64]),
        "orc_inst_free": (None, [P]),
        "orc_iso_apply": (None, [I32, P, P, P, I32, P]),
        "orc_iso_transform_aabb": (None, [I32, P, P, P, P]),
        "orc_ray_intersect": (I32, [I32, P, P, F32, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a, shape_last=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a


# --------------------------------------------------------------------------
# SplitMix64 as the reference tests draw it (sq/tests/conformance.rs:9-25)
# --------------------------------------------------------------------------

def rng_range(seed: int, lo: float, hi: float, n: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    lib().orc_rng_range(seed, lo, hi, n, _p(out))
    return out


def rng_u64(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.uint64)
    lib().orc_rng_u64(seed, n, _p(out))
    return out


def normalize(dirs: np.ndarray) -> np.ndarray:
    """Ray::new normalisation (ray.rs:18-23) of [n, D] directions."""
    d = np.array(dirs, dtype=np.float32, copy=True, order="C")
    lib().orc_normalize(d.shape[-1], _p(d), d.shape[0])
    return d


# --------------------------------------------------------------------------
# Single-level scene
# --------------------------------------------------------------------------

class Scene:
    """A built-in provider plus its reference-built Bvh (bvh.rs:31-37)."""

    def __init__(self, dim: int, kind: int, prims: np.ndarray, layers=None, build_tree: bool = True,
                 guarded: bool = False):
        """guarded: the engine's SQ_HITS_GUARDED restated (SPHERE and TRI hits
        kept only inside their own box's slab interval); default the plain
        providers, the reference's results."""
        self.dim, self.kind, self.guarded = dim, kind, guarded
        self.prims = np.ascontiguousarray(prims, dtype=np.float32)
        self.n = self.prims.shape[0]
        self._g = lib().orc_geom_new(dim, kind, _p(self.prims), self.n)
        if guarded:
            lib().orc_geom_set_guarded(self._g, 1)
        self.layers = None
        if layers is not None:
            self.set_layers(layers)
        self._b = lib().orc_bvh_build(self._g) if build_tree else None

    @classmethod
    def consumer(cls, dim: int, boxes: np.ndarray, prims: np.ndarray, narrow_addr: int, crossings_addr: int = 0,
                 cast_addr: int = 0, overlap_addr: int = 0, guarded: bool = False) -> "Scene":
        """A consumer provider: world boxes [n][2*dim] and consumer floats
        [n][pf], tested by the C function at narrow_addr (signature
        int(const float* prim, const float* origin, const float* dir, float
        max_toi, float* toi)); crossings through the C function at
        crossings_addr (signature void(prim, origin, dir, max_toi,
        void (*emit)(void*, float), void*)), or the nearest hit when 0; shape
        casts through the C function at cast_addr (int(prim, origin, dir,
        probe_min, probe_max, max_toi, float* toi); none when 0) and overlap
        through the one at overlap_addr (int(prim, region_min, region_max)
        after the box test; the box test alone when 0).  guarded: each hit,
        crossing and cast contact kept only inside its primitive's box slab
        interval (probe-grown for casts), the engine's SQ_HITS_GUARDED."""
        boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 2 * dim)
        prims = np.ascontiguousarray(prims, dtype=np.float32).reshape(boxes.shape[0], -1)
        obj = cls.__new__(cls)
        obj.dim, obj.kind, obj.guarded = dim, KIND_CONSUMER, guarded
        obj.prims = np.ascontiguousarray(np.concatenate([boxes, prims], 1), dtype=np.float32)
        obj.n = obj.prims.shape[0]
        obj._g = lib().orc_geom_new_consumer(dim, _p(obj.prims), obj.n, prims.shape[1], int(narrow_addr))
        if guarded:
            lib().orc_geom_set_guarded(obj._g, 1)
        if crossings_addr:
            lib().orc_geom_set_consumer_crossings(obj._g, int(crossings_addr))
        if cast_addr or overlap_addr:
            lib().orc_geom_set_consumer_cast_overlap(obj._g, int(cast_addr), int(overlap_addr))
        obj.layers = None
        obj._b = lib().orc_bvh_build(obj._g)
        return obj

    def set_layers(self, layers):
        self.layers = None if layers is None else np.ascontiguousarray(layers, dtype=np.uint64)
        lib().orc_geom_set_layers(self._g, _p(self.layers))

    def set_prims(self, prims):
        self.prims = np.ascontiguousarray(prims, dtype=np.float32)
        lib().orc_geom_set_prims(self._g, _p(self.prims))

    def refit(self):
        lib().orc_bvh_refit(self._b, self._g)

    def rebuild(self):
        if self._b:
            lib().orc_bvh_free(self._b)
        self._b = lib().orc_bvh_build(self._g)

    def world_aabbs(self):
        out = np.empty((self.n, 2 * self.dim), np.float32)
        lib().orc_geom_world_aabbs(self._g, _p(out))
        return out

    def export(self):
        """(bounds[nn,2D], child_a, child_b, prim_count, prim_indices) as bvh.rs stores them."""
        nn = lib().orc_bvh_node_count(self._b)
        bounds = np.empty((nn, 2 * self.dim), np.float32)
        ca = np.empty(nn, np.uint32); cb = np.empty(nn, np.uint32); pc = np.empty(nn, np.uint32)
        idx = np.empty(self.n, np.uint32)
        lib().orc_bvh_export(self._b, _p(bounds), _p(ca), _p(cb), _p(pc), _p(idx))
        return bounds, ca, cb, pc, idx

    def tree_stats(self):
        out = np.zeros(5, np.uint64)
        q = np.zeros(1, np.float32)
        lib().orc_bvh_stats(self._b, _p(out), _p(q))
        return dict(node_count=int(out[0]), leaf_count=int(out[1]), prim_count=int(out[2]),
                    max_depth=int(out[3]), bytes=int(out[4]), quality=float(q[0]))

    def raycast_nearest(self, rays, max_toi, mask=ALL_LAYERS, exclude=-1, threads=1, naive=False, stats=False):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 2 * self.dim)
        n = rays.shape[0]
        leaf = np.empty(n, np.int64); t = np.empty(n, np.float32)
        point = np.empty((n, self.dim), np.float32); normal = np.empty((n, self.dim), np.float32)
        sub = np.empty(n, np.int64)
        st = np.zeros((n, 4), np.uint64) if stats else None
        lib().orc_raycast_nearest_batch(self._b, self._g, _p(rays), n, max_toi, mask, exclude, threads,
                                        1 if naive else 0, _p(leaf), _p(t), _p(point), _p(normal), _p(sub), _p(st))
        out = dict(leaf=leaf, t=t, point=point, normal=normal, sub=sub)
        if stats:
            out["stats"] = st
        return out

    def raycast_all(self, rays, max_toi, mask=ALL_LAYERS, exclude=-1, naive=False, crossings=False):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 2 * self.dim)
        n = rays.shape[0]
        off = np.empty(n + 1, np.uint64)
        L = lib()
        total = L.orc_raycast_all_batch(self._b, self._g, _p(rays), n, max_toi, mask, exclude, int(naive), int(crossings),
                                        _p(off), None, None, 0)
        leaf = np.empty(total, np.int64); t = np.empty(total, np.float32)
        L.orc_raycast_all_batch(self._b, self._g, _p(rays), n, max_toi, mask, exclude, int(naive), int(crossings),
                                _p(off), _p(leaf), _p(t), total)
        return off, leaf, t

    def raycast_all_stats(self, rays, max_toi, mask=ALL_LAYERS, exclude=-1):
        """Bvh::raycast_all_stats' counters per ray (bvh.rs:206-215): uint64 [n][4]."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 2 * self.dim)
        st = np.zeros((rays.shape[0], 4), np.uint64)
        lib().orc_raycast_all_stats(self._b, self._g, _p(rays), rays.shape[0], max_toi, mask, exclude, _p(st))
        return st

    def overlap(self, regions, mask=ALL_LAYERS, exclude=-1, naive=False):
        regions = np.ascontiguousarray(regions, dtype=np.float32).reshape(-1, 2 * self.dim)
        n = regions.shape[0]
        off = np.empty(n + 1, np.uint64)
        L = lib()
        total = L.orc_overlap_batch(self._b, self._g, _p(regions), n, mask, exclude, int(naive), _p(off), None, 0)
        leaf = np.empty(total, np.int64)
        L.orc_overlap_batch(self._b, self._g, _p(regions), n, mask, exclude, int(naive), _p(off), _p(leaf), total)
        return off, leaf

    def shapecast_nearest(self, casts, max_toi, mask=ALL_LAYERS, exclude=-1, naive=False):
        casts = np.ascontiguousarray(casts, dtype=np.float32).reshape(-1, 4 * self.dim)
        n = casts.shape[0]
        mt = np.ascontiguousarray(np.broadcast_to(np.float32(max_toi) if np.isscalar(max_toi) else max_toi, (n,)), dtype=np.float32)
        leaf = np.empty(n, np.int64); t = np.empty(n, np.float32)
        point = np.empty((n, self.dim), np.float32); normal = np.empty((n, self.dim), np.float32)
        lib().orc_shapecast_nearest_batch(self._b, self._g, _p(casts), n, _p(mt), mask, exclude, int(naive),
                                          _p(leaf), _p(t), _p(point), _p(normal))
        return dict(leaf=leaf, t=t, point=point, normal=normal)

    def shapecast_all(self, casts, max_toi, mask=ALL_LAYERS, exclude=-1, naive=False):
        casts = np.ascontiguousarray(casts, dtype=np.float32).reshape(-1, 4 * self.dim)
        n = casts.shape[0]
        mt = np.ascontiguousarray(np.broadcast_to(np.float32(max_toi) if np.isscalar(max_toi) else max_toi, (n,)), dtype=np.float32)
        off = np.empty(n + 1, np.uint64)
        L = lib()
        total = L.orc_shapecast_all_batch(self._b, self._g, _p(casts), n, _p(mt), mask, exclude, int(naive), _p(off), None, None, 0)
        leaf = np.empty(total, np.int64); t = np.empty(total, np.float32)
        L.orc_shapecast_all_batch(self._b, self._g, _p(casts), n, _p(mt), mask, exclude, int(naive), _p(off), _p(leaf), _p(t), total)
        return off, leaf, t

    def required_work(self, rays, t_limit, threads=1):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 2 * self.dim)
        n = rays.shape[0]
        tl = np.ascontiguousarray(t_limit, dtype=np.float32)
        nodes = np.empty(n, np.uint64); prims = np.empty(n, np.uint64)
        lib().orc_required_work(self._b, _p(rays), n, _p(tl), _p(nodes), _p(prims), threads)
        return nodes, prims

    def __del__(self):
        try:
            L = lib()
            if getattr(self, "_b", None):
                L.orc_bvh_free(self._b)
            if getattr(self, "_g", None):
                L.orc_geom_free(self._g)
        except Exception:
            pass


# --------------------------------------------------------------------------
# Two-level scene (tlas.rs)
# --------------------------------------------------------------------------

F64_GUARDED, F64_BVH = 0x100, 0x200   # flags of the f64 oracle's `kind` argument (sq_oracle64.cpp)


class Scene64:
    """The reference's `f64` feature (sq/src/maths/point.rs:14-18) over BOX or
    SPHERE primitives in double: the reference's Bvh walks over its f64 tree
    (bvh.rs:51-456), or with naive=True the naive scans of
    sq/src/query/naive.rs; guarded=True restates the engine's SQ_HITS_GUARDED
    leaf tests (a sphere hit kept inside its own box), under which the two
    agree exactly (sq/src/lib.rs:15-17).  Also the f64 tree's stats."""

    def __init__(self, dim: int, kind: int, prims, layers=None, guarded: bool = False, naive: bool = False):
        self.dim, self.kind = dim, kind
        self._kf = kind | (F64_GUARDED if guarded else 0) | (0 if naive else F64_BVH)
        self.prims = np.ascontiguousarray(prims, np.float64).reshape(-1, 2 * dim if kind == BOX else dim + 1)
        self.n = self.prims.shape[0]
        self.layers = None if layers is None else np.ascontiguousarray(layers, np.uint64)

    def _q(self, q):
        return np.ascontiguousarray(q, np.float64).reshape(-1, 2 * self.dim)

    @staticmethod
    def _limits(max_toi, n):
        """A scalar max_toi, or one per query (returned as the array)."""
        if np.ndim(max_toi) == 0:
            return float(max_toi), None
        a = np.ascontiguousarray(max_toi, np.float64).reshape(n)
        return 0.0, a

    def raycast_nearest(self, rays, max_toi, mask=ALL_LAYERS, exclude=-1, threads=8):
        r = self._q(rays)
        n = r.shape[0]
        m, per = self._limits(max_toi, n)
        leaf = np.empty(n, np.int64); t = np.empty(n); pt = np.empty((n, self.dim)); nm = np.empty((n, self.dim))
        lib().orc64_raycast_nearest(self.dim, self._kf, _p(self.prims), self.n, _p(self.layers), mask, exclude, _p(r), n,
                                    m, _p(per), threads, _p(leaf), _p(t), _p(pt), _p(nm))
        return dict(leaf=leaf, t=t, point=pt, normal=nm)

    def _collect(self, fn, q, has_t):
        n = q.shape[0]
        off = np.empty(n + 1, np.uint64)
        tot = fn(off, None, None, 0)
        leaf = np.empty(max(tot, 1), np.int64); t = np.empty(max(tot, 1)) if has_t else None
        fn(off, leaf, t, tot)
        return off, leaf[:tot], (t[:tot] if has_t else None)

    def raycast_all(self, rays, max_toi, mask=ALL_LAYERS, exclude=-1, threads=8):
        r = self._q(rays)
        fn = lambda off, leaf, t, cap: lib().orc64_raycast_all(self.dim, self._kf, _p(self.prims), self.n, _p(self.layers),
                                                              mask, exclude, _p(r), r.shape[0], float(max_toi), threads,
                                                              _p(off), _p(leaf), _p(t), cap)
        return self._collect(fn, r, True)

    def overlap(self, regions, mask=ALL_LAYERS, exclude=-1, threads=8):
        r = self._q(regions)
        fn = lambda off, leaf, t, cap: lib().orc64_overlap(self.dim, self._kf, _p(self.prims), self.n, _p(self.layers),
                                                          mask, exclude, _p(r), r.shape[0], threads, _p(off), _p(leaf), cap)
        off, leaf, _ = self._collect(fn, r, False)
        return off, leaf

    def shapecast_nearest(self, casts, max_toi, mask=ALL_LAYERS, exclude=-1, threads=8):
        c = np.ascontiguousarray(casts, np.float64).reshape(-1, 4 * self.dim)
        n = c.shape[0]
        m, per = self._limits(max_toi, n)
        leaf = np.empty(n, np.int64); t = np.empty(n); pt = np.empty((n, self.dim))
        lib().orc64_shapecast_nearest(self.dim, self._kf, _p(self.prims), self.n, _p(self.layers), mask, exclude, _p(c), n,
                                      m, _p(per), threads, _p(leaf), _p(t), _p(pt))
        return dict(leaf=leaf, t=t, point=pt)

    def shapecast_all(self, casts, max_toi, mask=ALL_LAYERS, exclude=-1, threads=8):
        c = np.ascontiguousarray(casts, np.float64).reshape(-1, 4 * self.dim)
        m, per = self._limits(max_toi, c.shape[0])
        fn = lambda off, leaf, t, cap: lib().orc64_shapecast_all(self.dim, self._kf, _p(self.prims), self.n,
                                                                _p(self.layers), mask, exclude, _p(c), c.shape[0],
                                                                m, _p(per), threads, _p(off), _p(leaf), _p(t), cap)
        return self._collect(fn, c, True)

    def bvh_raycast_nearest(self, rays, max_toi, threads=1):
        """Bvh::raycast_nearest in reference order over the f64 tree, as
        raycast_nearest_batch_parallel (contiguous chunks); with the node boxes
        and primitives tested, summed over the rays."""
        r = self._q(rays)
        n = r.shape[0]
        leaf = np.empty(n, np.int64); t = np.empty(n); counts = np.zeros(2, np.uint64)
        lib().orc64_bvh_raycast_nearest(self.dim, self._kf, _p(self.prims), self.n, _p(r), n, float(max_toi), threads,
                                        _p(leaf), _p(t), _p(counts))
        return dict(leaf=leaf, t=t, nodes_tested=int(counts[0]), prims_tested=int(counts[1]))

    def bvh_raycast_nearest_refit(self, built_from, rays, max_toi, threads=1):
        """bvh_raycast_nearest over the tree built on the primitives
        `built_from` and refit to this scene's (Bvh::refit, bvh.rs:45-48:
        the build's topology, fresh bounds); with the refit tree's interior
        surface sum (tree.rs:104-110) as `quality`."""
        r = self._q(rays)
        n = r.shape[0]
        b0 = np.ascontiguousarray(built_from, np.float64).reshape(self.prims.shape)
        leaf = np.empty(n, np.int64); t = np.empty(n); counts = np.zeros(2, np.uint64); q = np.zeros(1)
        lib().orc64_bvh_refit_raycast_nearest(self.dim, self._kf, _p(b0), _p(self.prims), self.n, _p(r), n,
                                              float(max_toi), threads, _p(leaf), _p(t), _p(counts), _p(q))
        return dict(leaf=leaf, t=t, quality=float(q[0]))

    def tree_stats(self):
        out = np.zeros(5, np.uint64); q = np.zeros(1)
        lib().orc64_tree_stats(self.dim, self.kind, _p(self.prims), self.n, _p(out), _p(q))
        return dict(node_count=int(out[0]), leaf_count=int(out[1]), prim_count=int(out[2]), max_depth=int(out[3]),
                    bytes=int(out[4]), quality=float(q[0]))


class Consumer64:
    """A consumer provider under the `f64` feature: world boxes [n][2*dim] and
    the consumer's doubles [n][pd], its test_ray a C function at narrow_addr
    (int f(const double* prim, const double* origin, const double* dir,
    double max_toi, double* toi), the consumer's source built for the CPU).
    Nearest hits walk the reference's Bvh (bvh.rs:51-124, LIFO) or, with
    naive=True, scan every primitive (naive.rs); all hits as raycast_all
    (bvh.rs:196-259).  A hit counts in [0, max_toi] (geometry.rs:37-39);
    guarded=True keeps it only inside its box's slab interval (the engine's
    SQ_HITS_GUARDED)."""

    def __init__(self, dim: int, boxes, prims, narrow_addr: int, guarded: bool = False, naive: bool = False):
        self.dim = dim
        bx = np.ascontiguousarray(boxes, np.float64).reshape(-1, 2 * dim)
        pr = np.ascontiguousarray(prims, np.float64).reshape(bx.shape[0], -1)
        self.pd = pr.shape[1]
        self.records = np.ascontiguousarray(np.concatenate([bx, pr], 1))
        self.n = bx.shape[0]
        self.addr = int(narrow_addr)
        self._flags = (F64_GUARDED if guarded else 0) | (0 if naive else F64_BVH)

    def _q(self, q):
        return np.ascontiguousarray(q, np.float64).reshape(-1, 2 * self.dim)

    def raycast_nearest(self, rays, max_toi, threads=8):
        r = self._q(rays)
        n = r.shape[0]
        leaf = np.empty(n, np.int64); t = np.empty(n)
        lib().orc64_consumer_query(self.dim, self._flags, _p(self.records), self.n, self.pd, self.addr, _p(r), n,
                                   float(max_toi), 0, threads, _p(leaf), _p(t), None, 0)
        return dict(leaf=leaf, t=t)

    def raycast_all(self, rays, max_toi, threads=8):
        r = self._q(rays)
        n = r.shape[0]
        off = np.empty(n + 1, np.uint64)
        call = lambda leaf, t, cap: lib().orc64_consumer_query(self.dim, self._flags, _p(self.records), self.n, self.pd,
                                                               self.addr, _p(r), n, float(max_toi), 1, threads, _p(leaf),
                                                               _p(t), _p(off), cap)
        tot = call(None, None, 0)
        leaf = np.empty(max(tot, 1), np.int64); t = np.empty(max(tot, 1))
        call(leaf, t, tot)
        return off, leaf[:tot], t[:tot]

    def raycast_crossings(self, rays, max_toi, crossings_addr, threads=8):
        """Bvh::raycast_crossings (bvh.rs:264-328): every crossing the
        consumer's crossings function (C address) reports, per ray sorted by
        (toi, leaf)."""
        r = self._q(rays)
        n = r.shape[0]
        off = np.empty(n + 1, np.uint64)
        call = lambda leaf, t, cap: lib().orc64_consumer_crossings(self.dim, self._flags, _p(self.records), self.n, self.pd,
                                                                   self.addr, int(crossings_addr), _p(r), n, float(max_toi),
                                                                   threads, _p(leaf), _p(t), _p(off), cap)
        tot = call(None, None, 0)
        leaf = np.empty(max(tot, 1), np.int64); t = np.empty(max(tot, 1))
        call(leaf, t, tot)
        return off, leaf[:tot], t[:tot]


class BoxInstances64:
    """Box instances under the `f64` feature (two_level.rs:70-167 in double):
    the Tlas walk of tlas.rs:104-154 in the reference's order over the
    median-split tree of the world boxes (naive=True: the naive instance scan
    of two_level.rs:126-167), with guarded=True the engine's world-box guard;
    and the world boxes."""

    def __init__(self, dim, rot, trans, halves, guarded: bool = False, naive: bool = False):
        self.dim = dim
        self._flags = (F64_GUARDED if guarded else 0) | (0 if naive else F64_BVH)
        self.rot = np.ascontiguousarray(rot, np.float64).reshape(-1, dim, dim)
        self.trans = np.ascontiguousarray(trans, np.float64).reshape(-1, dim)
        self.halves = np.ascontiguousarray(halves, np.float64).reshape(-1, dim)
        self.n = self.rot.shape[0]

    def raycast_nearest(self, rays, max_toi, threads=8, world=False):
        r = np.ascontiguousarray(rays, np.float64).reshape(-1, 2 * self.dim)
        nr = r.shape[0]
        leaf = np.empty(nr, np.int64); t = np.empty(nr)
        w = np.empty((self.n, 2 * self.dim)) if world else None
        lib().orc64_tlas_box_nearest(self.dim, self.n, _p(self.rot), _p(self.trans), _p(self.halves), _p(r), nr,
                                     float(max_toi), threads, self._flags, _p(leaf), _p(t), _p(w))
        out = dict(leaf=leaf, t=t)
        if world: out["world"] = w
        return out


class Instances64:
    """The Tlas queries under the `f64` feature over box instances (kind 0,
    two_level.rs:70-113: halves [n][D]) or ball instances (kind 1,
    shape_overlap.rs:288-345: radii [n]): the Tlas walks of tlas.rs in the
    reference's order over the median-split tree of the world boxes
    (naive=True: the instance scans of two_level.rs:134-174 and naive.rs's
    rules), guarded=True the engine's world-box guard.  raycast_all is also
    Tlas::raycast_crossings for these kinds (one crossing per instance)."""

    def __init__(self, dim, kind, rot, trans, half, guarded: bool = False, naive: bool = False, built_from=None):
        """built_from: (rot, trans, half) of the instances the tree was built
        over; the walks then run over that tree refit to these instances'
        world boxes (Tlas::refit, tlas.rs:49-54)."""
        self.dim, self.kind = dim, kind
        self._flags = (F64_GUARDED if guarded else 0) | (0 if naive else F64_BVH)
        arrs = lambda r, t, h: (np.ascontiguousarray(r, np.float64).reshape(-1, dim, dim),
                                np.ascontiguousarray(t, np.float64).reshape(-1, dim),
                                np.ascontiguousarray(h, np.float64).reshape(-1, dim) if kind == 0 else
                                np.ascontiguousarray(h, np.float64).reshape(-1))
        self.rot, self.trans, self.half = arrs(rot, trans, half)
        self.built_from = None if built_from is None else arrs(*built_from)
        self.n = self.rot.shape[0]

    def _run(self, q, qs, max_toi, mode, off, leaf, t, cap, threads, quality=None):
        if self.built_from is None and quality is None:
            return lib().orc64_tlas_query(self.dim, self.kind, self.n, _p(self.rot), _p(self.trans), _p(self.half),
                                          _p(q), q.shape[0], float(max_toi), mode, threads, self._flags, _p(off),
                                          _p(leaf), _p(t), cap)
        r0, t0, h0 = self.built_from if self.built_from is not None else (self.rot, self.trans, self.half)
        return lib().orc64_tlas_refit_query(self.dim, self.kind, self.n, _p(r0), _p(t0), _p(h0), _p(self.rot),
                                            _p(self.trans), _p(self.half), _p(q), q.shape[0], float(max_toi), mode,
                                            threads, self._flags, _p(off), _p(leaf), _p(t), cap, _p(quality))

    def quality(self):
        """Interior surface sum of the (refit) tree, tree.rs:104-110."""
        q = np.zeros(1)
        self._run(np.zeros((0, 2 * self.dim)), 2 * self.dim, 1.0, 0, None, np.zeros(0, np.int64), np.zeros(0), 0, 1,
                  quality=q)
        return float(q[0])

    def needs_rebuild(self, factor, baseline):
        """Tlas::needs_rebuild (tlas.rs:58-60) against the build's quality."""
        return baseline > 0.0 and self.quality() > factor * baseline

    def _nearest(self, q, qs, max_toi, mode, threads):
        q = np.ascontiguousarray(q, np.float64).reshape(-1, qs)
        leaf = np.empty(q.shape[0], np.int64); t = np.empty(q.shape[0])
        self._run(q, qs, max_toi, mode, None, leaf, t, 0, threads)
        return dict(leaf=leaf, t=t)

    def _list(self, q, qs, max_toi, mode, threads):
        q = np.ascontiguousarray(q, np.float64).reshape(-1, qs)
        off = np.empty(q.shape[0] + 1, np.uint64)
        tot = self._run(q, qs, max_toi, mode, off, None, None, 0, threads)
        leaf = np.empty(max(tot, 1), np.int64); t = np.empty(max(tot, 1))
        self._run(q, qs, max_toi, mode, off, leaf, t, tot, threads)
        return off, leaf[:tot], t[:tot]

    def raycast_nearest(self, rays, max_toi, threads=8):
        return self._nearest(rays, 2 * self.dim, max_toi, 0, threads)

    def raycast_all(self, rays, max_toi, threads=8):
        return self._list(rays, 2 * self.dim, max_toi, 1, threads)

    def overlap(self, regions, threads=8):
        off, leaf, _ = self._list(regions, 2 * self.dim, 0.0, 2, threads)
        return off, leaf

    def shapecast_nearest(self, casts, max_toi, threads=8):
        return self._nearest(casts, 4 * self.dim, max_toi, 3, threads)

    def shapecast_all(self, casts, max_toi, threads=8):
        return self._list(casts, 4 * self.dim, max_toi, 4, threads)


class InstanceScene:
    def __init__(self, dim, kind, rot, trans, halves=None, mesh_tris=None, mesh_of=None, scale=None, guarded=False):
        """guarded: the engine's SQ_HITS_GUARDED restated (instance hits kept
        inside their world box's slab interval, mesh triangles guarded);
        default plain, the reference's Tlas results."""
        self.dim, self.kind = dim, kind
        self.rot = np.ascontiguousarray(rot, dtype=np.float32).reshape(-1, dim, dim)
        self.trans = np.ascontiguousarray(trans, dtype=np.float32).reshape(-1, dim)
        self.n = self.rot.shape[0]
        # INST_BALL: halves holds one radius per instance
        hw = 1 if kind == INST_BALL else dim
        self.halves = None if halves is None else np.ascontiguousarray(halves, dtype=np.float32).reshape(-1, hw)
        if kind == INST_MESH:
            off = np.zeros(len(mesh_tris) + 1, np.uint64)
            for i, m in enumerate(mesh_tris):
                off[i + 1] = off[i] + len(m)
            self.mesh_off = off
            self.tris = np.ascontiguousarray(np.concatenate([np.asarray(m, np.float32).reshape(-1, 9) for m in mesh_tris]), dtype=np.float32)
            self.mesh_of = np.ascontiguousarray(mesh_of, dtype=np.uint32)
            self.scale = np.ascontiguousarray(scale, dtype=np.float32)
            nm = len(mesh_tris)
        else:
            self.mesh_off = self.tris = self.mesh_of = self.scale = None
            nm = 0
        self._h = lib().orc_inst_new(dim, kind, self.n, _p(self.rot), _p(self.trans), _p(self.halves), nm,
                                     _p(self.mesh_off), _p(self.tris), _p(self.mesh_of), _p(self.scale))
        self.guarded = guarded or kind == INST_BOX_GUARDED
        if self.guarded:
            lib().orc_inst_set_guarded(self._h, 1)

    def set_transforms(self, rot, trans, halves=None):
        self.rot = np.ascontiguousarray(rot, dtype=np.float32).reshape(-1, self.dim, self.dim)
        self.trans = np.ascontiguousarray(trans, dtype=np.float32).reshape(-1, self.dim)
        if halves is not None:
            self.halves = np.ascontiguousarray(halves, dtype=np.float32).reshape(-1, 1 if self.kind == INST_BALL else self.dim)
        lib().orc_inst_set_transforms(self._h, _p(self.rot), _p(self.trans), _p(self.halves) if halves is not None else None)

    def refit(self):
        lib().orc_tlas_refit(self._h)

    def rebuild(self):
        lib().orc_tlas_rebuild(self._h)

    def needs_rebuild(self, factor):
        return bool(lib().orc_tlas_needs_rebuild(self._h, factor))

    def export(self):
        nn = lib().orc_tlas_node_count(self._h)
        bounds = np.empty((nn, 2 * self.dim), np.float32)
        ca = np.empty(nn, np.uint32); cb = np.empty(nn, np.uint32); pc = np.empty(nn, np.uint32)
        idx = np.empty(self.n, np.uint32)
        world = np.empty((self.n, 2 * self.dim), np.float32)
        lib().orc_tlas_export(self._h, _p(bounds), _p(ca), _p(cb), _p(pc), _p(idx), _p(world))
        return bounds, ca, cb, pc, idx, world

    def raycast_nearest(self, rays, max_toi, naive=False, threads=1):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 2 * self.dim)
        n = rays.shape[0]
        leaf = np.empty(n, np.int64); t = np.empty(n, np.float32)
        point = np.empty((n, self.dim), np.float32); normal = np.empty((n, self.dim), np.float32)
        sub = np.empty(n, np.int64)
        lib().orc_tlas_raycast_nearest_batch(self._h, _p(rays), n, max_toi, int(naive), threads,
                                             _p(leaf), _p(t), _p(point), _p(normal), _p(sub))
        return dict(leaf=leaf, t=t, point=point, normal=normal, sub=sub)

    def raycast_all(self, rays, max_toi):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 2 * self.dim)
        n = rays.shape[0]
        off = np.empty(n + 1, np.uint64)
        L = lib()
        total = L.orc_tlas_raycast_all_batch(self._h, _p(rays), n, max_toi, _p(off), None, None, None, 0)
        leaf = np.empty(total, np.int64); t = np.empty(total, np.float32); sub = np.empty(total, np.int64)
        L.orc_tlas_raycast_all_batch(self._h, _p(rays), n, max_toi, _p(off), _p(leaf), _p(t), _p(sub), total)
        return off, leaf, t, sub

    def shapecast_nearest(self, casts, max_toi):
        """Tlas::shapecast_nearest (tlas.rs:306-351) per cast [n][4D]: dict(leaf, t)."""
        casts = np.ascontiguousarray(casts, dtype=np.float32).reshape(-1, 4 * self.dim)
        n = casts.shape[0]
        leaf = np.empty(n, np.int64); t = np.empty(n, np.float32)
        lib().orc_tlas_shapecast_batch(self._h, _p(casts), n, max_toi, 0, None, _p(leaf), _p(t), n)
        return dict(leaf=leaf, t=t)

    def shapecast_all(self, casts, max_toi):
        """Tlas::shapecast_all (tlas.rs:353-391): CSR (offsets, leaf, t)."""
        casts = np.ascontiguousarray(casts, dtype=np.float32).reshape(-1, 4 * self.dim)
        n = casts.shape[0]
        off = np.empty(n + 1, np.uint64)
        L = lib()
        total = L.orc_tlas_shapecast_batch(self._h, _p(casts), n, max_toi, 1, _p(off), None, None, 0)
        leaf = np.empty(total, np.int64); t = np.empty(total, np.float32)
        L.orc_tlas_shapecast_batch(self._h, _p(casts), n, max_toi, 1, _p(off), _p(leaf), _p(t), total)
        return off, leaf, t

    def overlap(self, regions):
        regions = np.ascontiguousarray(regions, dtype=np.float32).reshape(-1, 2 * self.dim)
        n = regions.shape[0]
        off = np.empty(n + 1, np.uint64)
        L = lib()
        total = L.orc_tlas_overlap_batch(self._h, _p(regions), n, _p(off), None, 0)
        leaf = np.empty(total, np.int64)
        L.orc_tlas_overlap_batch(self._h, _p(regions), n, _p(off), _p(leaf), total)
        return off, leaf

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().orc_inst_free(self._h)
        except Exception:
            pass


def iso_apply(rot, trans, p, op):
    """op: 0 transform_point, 1 transform_vector, 2 inverse_transform_point, 3 inverse_transform_vector."""
    rot = np.ascontiguousarray(rot, np.float32); trans = np.ascontiguousarray(trans, np.float32)
    p = np.ascontiguousarray(p, np.float32)
    out = np.empty_like(p)
    lib().orc_iso_apply(p.shape[0], _p(rot), _p(trans), _p(p), op, _p(out))
    return out


def iso_transform_aabb(rot, trans, box):
    rot = np.ascontiguousarray(rot, np.float32); trans = np.ascontiguousarray(trans, np.float32)
    box = np.ascontiguousarray(box, np.float32)
    out = np.empty_like(box)
    lib().orc_iso_transform_aabb(box.shape[0] // 2, _p(rot), _p(trans), _p(box), _p(out))
    return out


def ray_intersect(box, ray, max_toi):
    box = np.ascontiguousarray(box, np.float32); ray = np.ascontiguousarray(ray, np.float32)
    a = np.zeros(1, np.float32); b = np.zeros(1, np.float32)
    hit = lib().orc_ray_intersect(box.shape[0] // 2, _p(box), _p(ray), max_toi, _p(a), _p(b))
    return (float(a[0]), float(b[0])) if hit else None
