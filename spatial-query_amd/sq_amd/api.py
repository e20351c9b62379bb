"""Host-side mirror of the reference API over the C ABI.

    Tree.build(dim, kind, prims)        Bvh::build(&G)            bvh.rs:31-37
    Tree.refit(prims)                   Bvh::refit(&G)            bvh.rs:45-48
    Tree.stats() / export()             Bvh::stats / nodes()      bvh.rs:465-489
    DeviceBvh(ctx, tree, prims)         GpuRaycaster::new         gpu/mod.rs:252-291
    DeviceBvh.raycast_nearest(...)      raycast_nearest_batch     bvh.rs:134-193, gpu/mod.rs:348-463
    DeviceBvh.raycast_any(...)          nearest(..).is_some()
    DeviceBvh.raycast_all(...)          raycast_all / _crossings  bvh.rs:196-328
    DeviceBvh.overlap(...)              overlap                   bvh.rs:418-456
    DeviceBvh.shapecast_nearest/all     shapecast_*               bvh.rs:331-414
    DeviceBvh.refit(prims)              GpuRaycaster::refit       gpu/mod.rs:326-341
"""
from __future__ import annotations

import atexit
import ctypes as C
import itertools
import sys
import weakref
from dataclasses import dataclass

import numpy as np

from .capi import BatchQuery, Filter, SqError, TreeStats, Tuning, check, lib, SQ_BATCH_BVH, SQ_BATCH_TLAS, SQ_ERR_CAPACITY

ALL_LAYERS = (1 << 64) - 1
SQ_HITS_REFERENCE, SQ_HITS_GUARDED = 0, 1   # sq_ctx_set_hit_mode (include/sq.h)

# Teardown order.  Engine objects release device memory and synchronise
# streams (the context's own, or torch's when the context shares it) when they
# close.  An object still alive when the interpreter shuts down (held by a
# module global, a saved traceback, a reference cycle) would otherwise be
# closed by the garbage collector during finalization, after torch's exit
# hooks have run and while the HIP runtime is being torn down; a C++
# exception thrown by the runtime there cannot be caught and ends the process
# with std::terminate.  So every object is tracked here; an exit hook, kept
# the newest one registered (so it runs before torch's, which atexit runs in
# reverse order of registration), closes the live ones newest first (trees
# before the contexts they hold), and __del__ does nothing once the
# interpreter is finalizing.
_live = weakref.WeakValueDictionary()
_serial = itertools.count()


def _close_all():
    for k in sorted(_live.keys(), reverse=True):
        obj = _live.get(k)
        if obj is None:
            continue
        try:
            obj.close()
        except Exception:
            pass


def _track(obj):
    """Registers an engine object for _close_all and keeps the exit hook last."""
    _live[next(_serial)] = obj
    atexit.unregister(_close_all)
    atexit.register(_close_all)
    return obj


def _finalize(obj):
    """The body of every __del__: close, unless the interpreter is shutting
    down (then the exit hook has closed it, or nothing safely can)."""
    if sys.is_finalizing():
        return
    try:
        obj.close()
    except Exception:
        pass


def prim_stride(dim: int, kind: int) -> int:
    return {0: 2 * dim, 1: dim + 1, 2: 9}[kind]


def _host(a, dtype=np.float32):
    return np.ascontiguousarray(a, dtype=dtype)


def _hp(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _dp(t, dtype: str | None = None, cols: int | None = None, n: int | None = None, per: int = 1):
    """Device pointer of a contiguous CUDA tensor (or None), after the checks
    that keep a kernel inside the buffer: the element type (`dtype`, a torch
    dtype name), for query arrays the row length (`cols`: a [n][cols] tensor),
    and for per-query inputs and outputs at least n * per elements.  A mismatch
    raises ValueError before anything reaches the device."""
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("query buffers must be contiguous GPU tensors")
    if dtype is not None and str(t.dtype) != "torch." + dtype:
        raise ValueError(f"expected a {dtype} tensor, got {t.dtype}")
    if cols is not None and (t.dim() != 2 or t.shape[1] != cols):
        raise ValueError(f"expected a [n][{cols}] tensor, got shape {tuple(t.shape)}")
    if n is not None and t.numel() < n * per:
        raise ValueError(f"expected at least {n * per} elements, got {t.numel()}")
    return C.c_void_p(t.data_ptr())


@dataclass
class QueryFilter:
    """QueryFilter (query/filter.rs:12-29): layer_mask, exclude (None = no exclusion)."""
    layer_mask: int = ALL_LAYERS
    exclude: int | None = None

    def c(self):
        return Filter(self.layer_mask & ALL_LAYERS, -1 if self.exclude is None else int(self.exclude))


class Context:
    """A device and its stream (GpuRaycaster::headless: raises SqError(NO_DEVICE) when absent)."""

    def __init__(self, device: int = 0, use_torch_stream: bool = True):
        ts = None
        if use_torch_stream:
            # torch sets up its device runtime first: a process whose HIP
            # runtime the engine started before torch's first device call
            # finds no devices in torch.  No device: sq_ctx_create reports it.
            import torch
            if torch.cuda.device_count() > device:
                ts = torch.cuda.current_stream(device).cuda_stream
        h = C.c_void_p()
        check(lib().sq_ctx_create(device, C.byref(h)))
        self._h = h
        self.device = device
        _track(self)
        if use_torch_stream:
            # Order engine work with the framework's work on the same stream.
            if ts is None:
                import torch
                ts = torch.cuda.current_stream(device).cuda_stream
            self.set_stream(ts)

    def set_stream(self, stream_ptr: int | None):
        """Enqueue on this hipStream_t (0 = the legacy default stream); None = the context's own."""
        if stream_ptr is None:
            check(lib().sq_ctx_reset_stream(self._h))
        else:
            check(lib().sq_ctx_set_stream(self._h, C.c_void_p(stream_ptr)))

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        check(lib().sq_ctx_synchronize(self._h))

    def stream_ptr(self) -> int:
        return lib().sq_ctx_stream(self._h) or 0

    def set_block(self, threads: int):
        check(lib().sq_ctx_set_block(self._h, threads))

    def set_tuning(self, block: int = 128, rays_per_lane: int = 16, lds_stack: int = 8, ray_order: int = 0,
                   descend_min: int = 24, split_depth: int = 0, wide: int = 2, packet: int = 0, lookahead: int = 1,
                   lookahead_pause: int = 40):
        check(lib().sq_ctx_set_tuning(self._h, C.byref(Tuning(block, rays_per_lane, lds_stack, ray_order, descend_min,
                                                              split_depth, wide, packet, lookahead, lookahead_pause))))

    def tuning(self) -> dict:
        t = Tuning()
        check(lib().sq_ctx_get_tuning(self._h, C.byref(t)))
        return dict(block=t.block, rays_per_lane=t.rays_per_lane, lds_stack=t.lds_stack, ray_order=t.ray_order,
                    descend_min=t.descend_min, split_depth=t.split_depth, wide=t.wide, packet=t.packet,
                    lookahead=t.lookahead, lookahead_pause=t.lookahead_pause)

    def set_hit_mode(self, mode: int):
        """SQ_HITS_REFERENCE (default: the reference's hits for every query) or
        SQ_HITS_GUARDED (SPHERE / TRIANGLE / instance hits kept only inside
        their own box's slab interval, walked near-first); see include/sq.h."""
        check(lib().sq_ctx_set_hit_mode(self._h, int(mode)))

    def hit_mode(self) -> int:
        return int(lib().sq_ctx_get_hit_mode(self._h))

    def last_kernel_ms(self) -> float:
        ms = C.c_float()
        check(lib().sq_ctx_last_kernel_ms(self._h, C.byref(ms)))
        return float(ms.value)

    def close(self):
        if getattr(self, "_h", None):
            lib().sq_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)


def _stats(s: TreeStats) -> dict:
    return dict(node_count=s.node_count, leaf_count=s.leaf_count, prim_count=s.prim_count,
                max_depth=s.max_depth, bytes=s.bytes, quality=float(s.quality))


class Tree:
    """The host tree in the reference layout (Bvh<D>): built by the reference
    median split, same topology node for node.  Needs no GPU."""

    def __init__(self, handle, dim: int, n_prims: int):
        self._h, self.dim, self.n_prims = handle, dim, n_prims
        _track(self)

    @classmethod
    def build(cls, dim: int, kind: int, prims) -> "Tree":
        p = _host(prims).reshape(-1, prim_stride(dim, kind))
        h = C.c_void_p()
        check(lib().sq_tree_build(dim, kind, _hp(p), p.shape[0], C.byref(h)))
        return cls(h, dim, p.shape[0])

    @classmethod
    def from_nodes(cls, dim, bounds, child_a, child_b, prim_count, prim_indices) -> "Tree":
        b = _host(bounds); ca = _host(child_a, np.uint32); cb = _host(child_b, np.uint32)
        pc = _host(prim_count, np.uint32); idx = _host(prim_indices, np.uint32)
        h = C.c_void_p()
        check(lib().sq_tree_from_nodes(dim, _hp(b), _hp(ca), _hp(cb), _hp(pc), len(ca), _hp(idx), len(idx), C.byref(h)))
        return cls(h, dim, len(idx))

    def refit(self, kind: int, prims):
        p = _host(prims)
        check(lib().sq_tree_refit(self._h, kind, _hp(p)))

    def stats(self) -> dict:
        s = TreeStats()
        check(lib().sq_tree_get_stats(self._h, C.byref(s)))
        return _stats(s)

    def export(self):
        nn = self.stats()["node_count"]
        bounds = np.empty((nn, 2 * self.dim), np.float32)
        ca = np.empty(nn, np.uint32); cb = np.empty(nn, np.uint32); pc = np.empty(nn, np.uint32)
        idx = np.empty(self.n_prims, np.uint32)
        check(lib().sq_tree_export(self._h, _hp(bounds), _hp(ca), _hp(cb), _hp(pc), _hp(idx)))
        return bounds, ca, cb, pc, idx

    def close(self):
        if getattr(self, "_h", None):
            lib().sq_tree_destroy(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)


class DeviceBvh:
    """A tree resident on the GPU with its leaf-ordered primitives."""

    def __init__(self, ctx: Context, tree: Tree, kind: int, prims):
        p = _host(prims).reshape(-1, prim_stride(tree.dim, kind))
        h = C.c_void_p()
        check(lib().sq_bvh_upload(ctx.handle, tree._h, kind, _hp(p), C.byref(h)))
        self._h, self.ctx, self.dim, self.kind, self.n_prims = h, ctx, tree.dim, kind, p.shape[0]
        _track(self)

    @classmethod
    def build(cls, ctx: Context, dim: int, kind: int, prims) -> "DeviceBvh":
        t = Tree.build(dim, kind, prims)
        try:
            return cls(ctx, t, kind, prims)
        finally:
            t.close()

    @classmethod
    def build_device(cls, ctx: Context, dim: int, kind: int, prims_dev) -> "DeviceBvh":
        """Bvh::build from primitives already on the GPU (a contiguous float32
        CUDA tensor in the kind's layout), built on the device: the same tree
        as build() node for node."""
        n = prims_dev.numel() // prim_stride(dim, kind)
        obj = cls.__new__(cls)
        h = C.c_void_p()
        check(lib().sq_bvh_build_device(ctx.handle, dim, kind, _dp(prims_dev, "float32") if n else None, n, C.byref(h)))
        obj._h, obj.ctx, obj.dim, obj.kind, obj.n_prims = h, ctx, dim, kind, n
        return _track(obj)

    def export(self):
        """(bounds, child_a, child_b, prim_count, prim_indices) of the device tree
        in the reference layout (Bvh::nodes / prim_indices)."""
        nn = self.stats()["node_count"]
        bounds = np.empty((nn, 2 * self.dim), np.float32)
        ca = np.empty(nn, np.uint32); cb = np.empty(nn, np.uint32); pc = np.empty(nn, np.uint32)
        idx = np.empty(self.n_prims, np.uint32)
        check(lib().sq_bvh_export(self._h, _hp(bounds), _hp(ca), _hp(cb), _hp(pc), _hp(idx)))
        return bounds, ca, cb, pc, idx

    def rebuild(self, prims_dev):
        """Build the tree again, on the device, from moved primitives (same count)."""
        check(lib().sq_bvh_rebuild(self._h, _dp(prims_dev, "float32", n=self.n_prims, per=self._stride())))

    def surface_sum(self) -> float:
        """Interior surface sum of the current device bounds (tree.rs:104-110), summed on the device."""
        v = C.c_double()
        check(lib().sq_bvh_surface_sum(self._h, C.byref(v)))
        return v.value

    def refit_auto(self, prims_dev, factor: float = 1.5) -> bool:
        """Refit to moved primitives, and rebuild on the device once the surface sum exceeds factor x the
        sum of the tree as last built (Tlas::needs_rebuild's test); True when it rebuilt."""
        r = C.c_int()
        check(lib().sq_bvh_refit_auto(self._h, _dp(prims_dev, "float32", n=self.n_prims, per=self._stride()),
                                      float(factor), C.byref(r)))
        return bool(r.value)

    # -- management ---------------------------------------------------------
    def set_layers(self, layers):
        l = None if layers is None else _host(layers, np.uint64)
        check(lib().sq_bvh_set_layers(self._h, _hp(l)))

    def stats(self) -> dict:
        s = TreeStats()
        check(lib().sq_bvh_get_stats(self._h, C.byref(s)))
        return _stats(s)

    def download_bounds(self):
        nn = self.stats()["node_count"]
        out = np.empty((nn, 2 * self.dim), np.float32)
        check(lib().sq_bvh_download_bounds(self._h, _hp(out)))
        return out

    def refit(self, prims_dev):
        check(lib().sq_bvh_refit(self._h, _dp(prims_dev, "float32", n=self.n_prims, per=self._stride())))

    def _stride(self) -> int:
        # consumer trees (kind 3) hold records the consumer laid out: not checked here
        return prim_stride(self.dim, self.kind) if self.kind in (0, 1, 2) else 0

    # -- queries --------------------------------------------------------------
    def _f(self, filt):
        return None if filt is None else C.byref(filt.c())

    def raycast_nearest(self, rays, max_toi, max_toi_per_ray=None, filt=None, point=False, normal=False, sub=False):
        import torch
        n = rays.shape[0]
        dev = rays.device
        leaf = torch.empty(n, dtype=torch.int32, device=dev)
        toi = torch.empty(n, dtype=torch.float32, device=dev)
        pt = torch.empty((n, self.dim), dtype=torch.float32, device=dev) if point else None
        nm = torch.empty((n, self.dim), dtype=torch.float32, device=dev) if normal else None
        sb = torch.empty(n, dtype=torch.int32, device=dev) if sub else None
        self.raycast_nearest_into(rays, max_toi, leaf, toi, max_toi_per_ray, filt, pt, nm, sb)
        out = dict(leaf=leaf, toi=toi)
        if point: out["point"] = pt
        if normal: out["normal"] = nm
        if sub: out["sub"] = sb
        return out

    def raycast_nearest_into(self, rays, max_toi, leaf, toi, max_toi_per_ray=None, filt=None, point=None, normal=None, sub=None):
        n = rays.shape[0]
        check(lib().sq_raycast_nearest(self._h, _dp(rays, "float32", 2 * self.dim), n, float(max_toi),
                                       _dp(max_toi_per_ray, "float32", n=n), self._f(filt), _dp(leaf, "int32", n=n),
                                       _dp(toi, "float32", n=n), _dp(point, "float32", n=n, per=self.dim),
                                       _dp(normal, "float32", n=n, per=self.dim), _dp(sub, "int32", n=n)))

    def raycast_nearest_noalloc_into(self, rays, max_toi, leaf, toi, max_toi_per_ray=None, filt=None, point=None,
                                     normal=None, sub=None):
        """sq_raycast_nearest_noalloc: raycast_nearest_into without any device
        allocation (the walk's counters and stack spill in the library's fixed
        per-device array); the caller provides every output."""
        n = rays.shape[0]
        check(lib().sq_raycast_nearest_noalloc(self._h, _dp(rays, "float32", 2 * self.dim), n, float(max_toi),
                                               _dp(max_toi_per_ray, "float32", n=n), self._f(filt),
                                               _dp(leaf, "int32", n=n), _dp(toi, "float32", n=n),
                                               _dp(point, "float32", n=n, per=self.dim),
                                               _dp(normal, "float32", n=n, per=self.dim), _dp(sub, "int32", n=n)))

    def raycast_any(self, rays, max_toi, max_toi_per_ray=None, filt=None, out=None):
        import torch
        if out is None:
            out = torch.empty(rays.shape[0], dtype=torch.uint8, device=rays.device)
        check(lib().sq_raycast_any(self._h, _dp(rays, "float32", 2 * self.dim), rays.shape[0], float(max_toi), _dp(max_toi_per_ray, "float32", n=rays.shape[0]),
                                   self._f(filt), _dp(out, "uint8", n=rays.shape[0])))
        return out

    def raycast_nearest_stats(self, rays, max_toi, max_toi_per_ray=None, filt=None):
        """Bvh::raycast_nearest_stats per ray: dict(leaf, toi, stats int32 [n][4] = nodes_visited, aabb_tests,
        narrow_tests, hits), counted on a walk in the reference's order over its hits (a diagnostic walk)."""
        import torch
        n, dev = rays.shape[0], rays.device
        leaf = torch.empty(n, dtype=torch.int32, device=dev)
        toi = torch.empty(n, dtype=torch.float32, device=dev)
        st = torch.empty((n, 4), dtype=torch.int32, device=dev)
        check(lib().sq_raycast_nearest_stats(self._h, _dp(rays, "float32", 2 * self.dim), n, float(max_toi),
                                             _dp(max_toi_per_ray, "float32", n=n), self._f(filt), _dp(leaf, "int32", n=n),
                                             _dp(toi, "float32", n=n), _dp(st, "int32", n=n, per=4)))
        return dict(leaf=leaf, toi=toi, stats=st)

    def raycast_all_stats(self, rays, max_toi, filt=None):
        """Bvh::raycast_all_stats' counters per ray (int32 [n][4]); the hit lists come from raycast_all."""
        import torch
        n = rays.shape[0]
        st = torch.empty((n, 4), dtype=torch.int32, device=rays.device)
        check(lib().sq_raycast_all_stats(self._h, _dp(rays, "float32", 2 * self.dim), n, float(max_toi), self._f(filt),
                                         _dp(st, "int32", n=n, per=4)))
        return st

    def _csr(self, fn, n, dev, has_t, capacity):
        import torch
        offsets = torch.empty(n + 1, dtype=torch.int64, device=dev)
        total = C.c_uint64()
        # capacity 0: start from the last call's total (or 4 entries per query),
        # so that a call whose result fits walks the queries once
        cap = capacity or max(getattr(self, "_csr_hint", 0), 4 * n)
        while True:
            leaf = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
            toi = torch.empty(max(cap, 1), dtype=torch.float32, device=dev) if has_t else None
            rc = fn(offsets, leaf, toi, cap, total)
            if rc == SQ_ERR_CAPACITY:
                cap = int(total.value)
                continue
            check(rc)
            break
        t = int(total.value)
        self._csr_hint = t
        return offsets, leaf[:t], (toi[:t] if has_t else None)

    def raycast_all(self, rays, max_toi, filt=None, crossings=False, capacity=0):
        def fn(off, leaf, toi, cap, total):
            return lib().sq_raycast_all(self._h, _dp(rays, "float32", 2 * self.dim), rays.shape[0], float(max_toi), self._f(filt), int(crossings),
                                        _dp(off, "int64"), _dp(leaf, "int32"), _dp(toi, "float32"), cap, C.byref(total))
        return self._csr(fn, rays.shape[0], rays.device, True, capacity)

    def overlap(self, regions, filt=None, capacity=0):
        def fn(off, leaf, toi, cap, total):
            return lib().sq_overlap(self._h, _dp(regions, "float32", 2 * self.dim), regions.shape[0], self._f(filt), _dp(off, "int64"), _dp(leaf, "int32"),
                                    cap, C.byref(total))
        off, leaf, _ = self._csr(fn, regions.shape[0], regions.device, False, capacity)
        return off, leaf

    def shapecast_nearest(self, casts, max_toi, max_toi_per_cast=None, filt=None, point=False):
        import torch
        n = casts.shape[0]
        dev = casts.device
        leaf = torch.empty(n, dtype=torch.int32, device=dev)
        toi = torch.empty(n, dtype=torch.float32, device=dev)
        pt = torch.empty((n, self.dim), dtype=torch.float32, device=dev) if point else None
        check(lib().sq_shapecast_nearest(self._h, _dp(casts, "float32", 4 * self.dim), n, float(max_toi), _dp(max_toi_per_cast, "float32", n=casts.shape[0]), self._f(filt),
                                         _dp(leaf, "int32"), _dp(toi, "float32"), _dp(pt, "float32")))
        out = dict(leaf=leaf, toi=toi)
        if point: out["point"] = pt
        return out

    def shapecast_all(self, casts, max_toi, max_toi_per_cast=None, filt=None, capacity=0):
        def fn(off, leaf, toi, cap, total):
            return lib().sq_shapecast_all(self._h, _dp(casts, "float32", 4 * self.dim), casts.shape[0], float(max_toi), _dp(max_toi_per_cast, "float32", n=casts.shape[0]),
                                          self._f(filt), _dp(off, "int64"), _dp(leaf, "int32"), _dp(toi, "float32"), cap, C.byref(total))
        return self._csr(fn, casts.shape[0], casts.device, True, capacity)

    def required_work(self, rays, t_limit):
        import torch
        n = rays.shape[0]
        nodes = torch.empty(n, dtype=torch.int32, device=rays.device)
        prims = torch.empty(n, dtype=torch.int32, device=rays.device)
        check(lib().sq_required_work(self._h, _dp(rays, "float32", 2 * self.dim), n, _dp(t_limit, "float32", n=n), _dp(nodes, "int32"), _dp(prims, "int32")))
        return nodes, prims

    def close(self):
        if getattr(self, "_h", None):
            lib().sq_bvh_destroy(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)


class DeviceBvh64:
    """The reference's `f64` feature (maths/point.rs:14-18): a tree over
    double-precision AABB or SPHERE primitives, built with the median split in
    double and walked on the GPU in double.  Queries take float64 CUDA tensors
    and answer as DeviceBvh does, with float64 toi, points and normals."""

    def __init__(self, ctx: Context, dim: int, kind: int, prims):
        p = _host(prims, np.float64).reshape(-1, prim_stride(dim, kind))
        h = C.c_void_p()
        check(lib().sq_bvh64_create(ctx.handle, dim, kind, _hp(p), p.shape[0], C.byref(h)))
        self._h, self.ctx, self.dim, self.kind, self.n_prims = h, ctx, dim, kind, p.shape[0]
        _track(self)

    def refit(self, prims):
        """Bvh::refit in double: the same number of primitives at new
        positions (host float64, the layout of the build)."""
        p = _host(prims, np.float64).reshape(-1, prim_stride(self.dim, self.kind))
        if p.shape[0] != self.n_prims:
            raise ValueError(f"refit takes {self.n_prims} primitives, got {p.shape[0]}")
        check(lib().sq_bvh64_refit(self._h, _hp(p)))

    def set_layers(self, layers):
        l = None if layers is None else _host(layers, np.uint64)
        check(lib().sq_bvh64_set_layers(self._h, _hp(l)))

    def stats(self) -> dict:
        s = TreeStats()
        check(lib().sq_bvh64_stats(self._h, C.byref(s)))
        return _stats(s)

    def _f(self, filt):
        return None if filt is None else C.byref(filt.c())

    def raycast_nearest(self, rays, max_toi, max_toi_per_ray=None, filt=None, point=False, normal=False):
        import torch
        n, dev = rays.shape[0], rays.device
        leaf = torch.empty(n, dtype=torch.int32, device=dev)
        toi = torch.empty(n, dtype=torch.float64, device=dev)
        pt = torch.empty((n, self.dim), dtype=torch.float64, device=dev) if point else None
        nm = torch.empty((n, self.dim), dtype=torch.float64, device=dev) if normal else None
        check(lib().sq_bvh64_raycast_nearest(self._h, _dp(rays, "float64", 2 * self.dim), n, float(max_toi), _dp(max_toi_per_ray, "float64", n=rays.shape[0]), self._f(filt),
                                             _dp(leaf, "int32"), _dp(toi, "float64"), _dp(pt, "float64"), _dp(nm, "float64")))
        out = dict(leaf=leaf, toi=toi)
        if point: out["point"] = pt
        if normal: out["normal"] = nm
        return out

    def raycast_any(self, rays, max_toi, max_toi_per_ray=None, filt=None):
        import torch
        out = torch.empty(rays.shape[0], dtype=torch.uint8, device=rays.device)
        check(lib().sq_bvh64_raycast_any(self._h, _dp(rays, "float64", 2 * self.dim), rays.shape[0], float(max_toi), _dp(max_toi_per_ray, "float64", n=rays.shape[0]),
                                         self._f(filt), _dp(out, "uint8", n=rays.shape[0])))
        return out

    def _csr(self, fn, n, dev, has_t, capacity):
        import torch
        offsets = torch.empty(n + 1, dtype=torch.int64, device=dev)
        total = C.c_uint64()
        cap = capacity or max(getattr(self, "_csr_hint", 0), 4 * n)
        while True:
            leaf = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
            toi = torch.empty(max(cap, 1), dtype=torch.float64, device=dev) if has_t else None
            rc = fn(offsets, leaf, toi, cap, total)
            if rc == SQ_ERR_CAPACITY:
                cap = int(total.value)
                continue
            check(rc)
            break
        t = int(total.value)
        self._csr_hint = t
        return offsets, leaf[:t], (toi[:t] if has_t else None)

    def raycast_all(self, rays, max_toi, filt=None, capacity=0):
        def fn(off, leaf, toi, cap, total):
            return lib().sq_bvh64_raycast_all(self._h, _dp(rays, "float64", 2 * self.dim), rays.shape[0], float(max_toi), self._f(filt),
                                              _dp(off, "int64"), _dp(leaf, "int32"), _dp(toi, "float64"), cap, C.byref(total))
        return self._csr(fn, rays.shape[0], rays.device, True, capacity)

    def shapecast_nearest(self, casts, max_toi, max_toi_per_cast=None, filt=None, point=False):
        import torch
        n, dev = casts.shape[0], casts.device
        leaf = torch.empty(n, dtype=torch.int32, device=dev)
        toi = torch.empty(n, dtype=torch.float64, device=dev)
        pt = torch.empty((n, self.dim), dtype=torch.float64, device=dev) if point else None
        check(lib().sq_bvh64_shapecast_nearest(self._h, _dp(casts, "float64", 4 * self.dim), n, float(max_toi), _dp(max_toi_per_cast, "float64", n=casts.shape[0]),
                                               self._f(filt), _dp(leaf, "int32"), _dp(toi, "float64"), _dp(pt, "float64")))
        out = dict(leaf=leaf, toi=toi)
        if point: out["point"] = pt
        return out

    def shapecast_all(self, casts, max_toi, max_toi_per_cast=None, filt=None, capacity=0):
        def fn(off, leaf, toi, cap, total):
            return lib().sq_bvh64_shapecast_all(self._h, _dp(casts, "float64", 4 * self.dim), casts.shape[0], float(max_toi), _dp(max_toi_per_cast, "float64", n=casts.shape[0]),
                                                self._f(filt), _dp(off, "int64"), _dp(leaf, "int32"), _dp(toi, "float64"), cap, C.byref(total))
        return self._csr(fn, casts.shape[0], casts.device, True, capacity)

    def overlap(self, regions, filt=None, capacity=0):
        def fn(off, leaf, toi, cap, total):
            return lib().sq_bvh64_overlap(self._h, _dp(regions, "float64", 2 * self.dim), regions.shape[0], self._f(filt), _dp(off, "int64"), _dp(leaf, "int32"),
                                          cap, C.byref(total))
        off, leaf, _ = self._csr(fn, regions.shape[0], regions.device, False, capacity)
        return off, leaf

    def close(self):
        if getattr(self, "_h", None):
            lib().sq_bvh64_destroy(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)


class Tlas64:
    """Tlas under the `f64` feature over box instances (kind 0: rotations
    [n][D][D], translations [n][D], half extents [n][D]) or ball instances
    (kind 2, D = 3: halves = radii [n]); host float64.  raycast_all is also
    Tlas::raycast_crossings for these kinds (one crossing per instance)."""

    def __init__(self, ctx: Context, dim: int, rot, trans, halves, kind: int = 0):
        r = _host(rot, np.float64).reshape(-1, dim, dim)
        t = _host(trans, np.float64).reshape(-1, dim)
        h = _host(halves, np.float64).reshape(-1) if kind == 2 else _host(halves, np.float64).reshape(-1, dim)
        if h.shape[0] != r.shape[0] or t.shape[0] != r.shape[0]:
            raise ValueError("rot, trans and halves must describe the same number of instances")
        hd = C.c_void_p()
        check(lib().sq_tlas64_create_kind(ctx.handle, dim, int(kind), r.shape[0], _hp(r), _hp(t), _hp(h), C.byref(hd)))
        self._h, self.ctx, self.dim, self.n, self.kind = hd, ctx, dim, r.shape[0], kind
        _track(self)

    def refit(self, rot, trans, halves=None):
        """Tlas::refit in double: new transforms (host float64, as at the
        build), and half extents / radii (None keeps them)."""
        r = _host(rot, np.float64).reshape(-1, self.dim, self.dim)
        t = _host(trans, np.float64).reshape(-1, self.dim)
        h = None if halves is None else (_host(halves, np.float64).reshape(-1) if self.kind == 2 else
                                         _host(halves, np.float64).reshape(-1, self.dim))
        if r.shape[0] != self.n or t.shape[0] != self.n or (h is not None and h.shape[0] != self.n):
            raise ValueError(f"refit takes the {self.n} instances of the build")
        check(lib().sq_tlas64_refit(self._h, _hp(r), _hp(t), _hp(h)))

    def needs_rebuild(self, factor: float) -> bool:
        out = C.c_int()
        check(lib().sq_tlas64_needs_rebuild(self._h, float(factor), C.byref(out)))
        return bool(out.value)

    _csr = DeviceBvh64._csr

    def raycast_all(self, rays, max_toi, capacity=0):
        def fn(off, leaf, toi, cap, total):
            return lib().sq_tlas64_raycast_all(self._h, _dp(rays, "float64", 2 * self.dim), rays.shape[0], float(max_toi),
                                               _dp(off, "int64"), _dp(leaf, "int32"), _dp(toi, "float64"), cap, C.byref(total))
        return self._csr(fn, rays.shape[0], rays.device, True, capacity)

    def overlap(self, regions, capacity=0):
        def fn(off, leaf, toi, cap, total):
            return lib().sq_tlas64_overlap(self._h, _dp(regions, "float64", 2 * self.dim), regions.shape[0], _dp(off, "int64"),
                                           _dp(leaf, "int32"), cap, C.byref(total))
        off, leaf, _ = self._csr(fn, regions.shape[0], regions.device, False, capacity)
        return off, leaf

    def shapecast_nearest(self, casts, max_toi, point=False):
        import torch
        n, dev = casts.shape[0], casts.device
        leaf = torch.empty(n, dtype=torch.int32, device=dev)
        toi = torch.empty(n, dtype=torch.float64, device=dev)
        pt = torch.empty((n, self.dim), dtype=torch.float64, device=dev) if point else None
        check(lib().sq_tlas64_shapecast_nearest(self._h, _dp(casts, "float64", 4 * self.dim), n, float(max_toi),
                                                _dp(leaf, "int32"), _dp(toi, "float64"), _dp(pt, "float64")))
        out = dict(leaf=leaf, toi=toi)
        if point: out["point"] = pt
        return out

    def shapecast_all(self, casts, max_toi, capacity=0):
        def fn(off, leaf, toi, cap, total):
            return lib().sq_tlas64_shapecast_all(self._h, _dp(casts, "float64", 4 * self.dim), casts.shape[0], float(max_toi),
                                                 _dp(off, "int64"), _dp(leaf, "int32"), _dp(toi, "float64"), cap, C.byref(total))
        return self._csr(fn, casts.shape[0], casts.device, True, capacity)

    def stats(self) -> dict:
        s = TreeStats()
        check(lib().sq_tlas64_stats(self._h, C.byref(s)))
        return _stats(s)

    def raycast_nearest(self, rays, max_toi, max_toi_per_ray=None, point=False):
        import torch
        n, dev = rays.shape[0], rays.device
        leaf = torch.empty(n, dtype=torch.int32, device=dev)
        toi = torch.empty(n, dtype=torch.float64, device=dev)
        pt = torch.empty((n, self.dim), dtype=torch.float64, device=dev) if point else None
        check(lib().sq_tlas64_raycast_nearest(self._h, _dp(rays, "float64", 2 * self.dim), n, float(max_toi), _dp(max_toi_per_ray, "float64", n=rays.shape[0]), _dp(leaf, "int32"),
                                              _dp(toi, "float64"), _dp(pt, "float64")))
        out = dict(leaf=leaf, toi=toi)
        if point: out["point"] = pt
        return out

    def raycast_any(self, rays, max_toi, max_toi_per_ray=None):
        import torch
        out = torch.empty(rays.shape[0], dtype=torch.uint8, device=rays.device)
        check(lib().sq_tlas64_raycast_any(self._h, _dp(rays, "float64", 2 * self.dim), rays.shape[0], float(max_toi), _dp(max_toi_per_ray, "float64", n=rays.shape[0]),
                                          _dp(out, "uint8", n=rays.shape[0])))
        return out

    def close(self):
        if getattr(self, "_h", None):
            lib().sq_tlas64_destroy(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)


class Tlas:
    """Tlas<D> over built-in instance kinds (accel/tlas.rs), resident on the GPU.

    kind 0 (box): halves[n, D] local boxes; kind 1 (mesh, D = 3): meshes is a
    list of TRIANGLE DeviceBvh, mesh_of[n], scale[n] uniform scales; kind 2
    (ball, D = 3): halves[n] radii of balls at the local origin."""

    BOX, MESH, BALL = 0, 1, 2

    def __init__(self, ctx: Context, dim: int, kind: int, rot, trans, halves=None, meshes=None, mesh_of=None, scale=None):
        rot = _host(rot).reshape(-1, dim, dim); trans = _host(trans).reshape(-1, dim)
        n = rot.shape[0]
        self._keep = []
        hv = _host(halves).reshape(-1, 1 if kind == self.BALL else dim) if halves is not None else None
        if kind == self.MESH:
            arr = (C.c_void_p * len(meshes))(*[m._h.value for m in meshes])
            mo = _host(mesh_of, np.uint32); sc = _host(scale)
            self._keep = [arr, mo, sc, list(meshes)]
            mesh_ptr, nm = C.cast(arr, C.c_void_p), len(meshes)
        else:
            mesh_ptr, nm, mo, sc = None, 0, None, None
        h = C.c_void_p()
        check(lib().sq_tlas_build(ctx.handle, dim, kind, n, _hp(rot), _hp(trans), _hp(hv), mesh_ptr, nm, _hp(mo), _hp(sc),
                                  C.byref(h)))
        self._h, self.ctx, self.dim, self.kind, self.n = h, ctx, dim, kind, n
        _track(self)

    def refit(self, rot, trans, halves=None):
        rot = _host(rot); trans = _host(trans)
        hv = _host(halves) if halves is not None else None
        check(lib().sq_tlas_refit(self._h, _hp(rot), _hp(trans), _hp(hv)))

    def needs_rebuild(self, factor: float) -> bool:
        out = C.c_int()
        check(lib().sq_tlas_needs_rebuild(self._h, factor, C.byref(out)))
        return bool(out.value)

    def stats(self) -> dict:
        s = TreeStats()
        check(lib().sq_tlas_get_stats(self._h, C.byref(s)))
        return _stats(s)

    def download_bounds(self):
        nn = self.stats()["node_count"]
        out = np.empty((nn, 2 * self.dim), np.float32)
        check(lib().sq_tlas_download_bounds(self._h, _hp(out)))
        return out

    def raycast_nearest(self, rays, max_toi, max_toi_per_ray=None, point=False, normal=False, sub=False):
        import torch
        n = rays.shape[0]
        dev = rays.device
        inst = torch.empty(n, dtype=torch.int32, device=dev)
        toi = torch.empty(n, dtype=torch.float32, device=dev)
        pt = torch.empty((n, self.dim), dtype=torch.float32, device=dev) if point else None
        nm = torch.empty((n, self.dim), dtype=torch.float32, device=dev) if normal else None
        sb = torch.empty(n, dtype=torch.int32, device=dev) if sub else None
        check(lib().sq_tlas_raycast_nearest(self._h, _dp(rays, "float32", 2 * self.dim), n, float(max_toi), _dp(max_toi_per_ray, "float32", n=rays.shape[0]), _dp(inst, "int32"),
                                            _dp(toi, "float32"), _dp(pt, "float32"), _dp(nm, "float32"), _dp(sb, "int32")))
        out = dict(leaf=inst, toi=toi)
        if point: out["point"] = pt
        if normal: out["normal"] = nm
        if sub: out["sub"] = sb
        return out

    def raycast_any(self, rays, max_toi, max_toi_per_ray=None):
        import torch
        out = torch.empty(rays.shape[0], dtype=torch.uint8, device=rays.device)
        check(lib().sq_tlas_raycast_any(self._h, _dp(rays, "float32", 2 * self.dim), rays.shape[0], float(max_toi), _dp(max_toi_per_ray, "float32", n=rays.shape[0]), _dp(out, "uint8", n=rays.shape[0])))
        return out

    def raycast_all(self, rays, max_toi, capacity=0):
        def fn(off, leaf, toi, cap, total):
            return lib().sq_tlas_raycast_all(self._h, _dp(rays, "float32", 2 * self.dim), rays.shape[0], float(max_toi), _dp(off, "int64"), _dp(leaf, "int32"),
                                             _dp(toi, "float32"), cap, C.byref(total))
        return DeviceBvh._csr(self, fn, rays.shape[0], rays.device, True, capacity)

    def overlap(self, regions, capacity=0):
        def fn(off, leaf, toi, cap, total):
            return lib().sq_tlas_overlap(self._h, _dp(regions, "float32", 2 * self.dim), regions.shape[0], _dp(off, "int64"), _dp(leaf, "int32"), cap, C.byref(total))
        off, leaf, _ = DeviceBvh._csr(self, fn, regions.shape[0], regions.device, False, capacity)
        return off, leaf

    def raycast_crossings(self, rays, max_toi, capacity=0):
        """Tlas::raycast_crossings (tlas.rs:227-302): the built-in instance kinds
        report one crossing per instance (test_ray_crossings_local's default), so
        this is raycast_all."""
        return self.raycast_all(rays, max_toi, capacity)

    def shapecast_nearest(self, casts, max_toi, point=False):
        """Tlas::shapecast_nearest, casts [n][4D] (origin, dir, probe box min, max): dict(leaf, toi[, point])."""
        import torch
        n = casts.shape[0]
        leaf = torch.empty(n, dtype=torch.int32, device=casts.device)
        toi = torch.empty(n, dtype=torch.float32, device=casts.device)
        pt = torch.empty((n, self.dim), dtype=torch.float32, device=casts.device) if point else None
        check(lib().sq_tlas_shapecast_nearest(self._h, _dp(casts, "float32", 4 * self.dim), n, float(max_toi), _dp(leaf, "int32"), _dp(toi, "float32"), _dp(pt, "float32")))
        out = dict(leaf=leaf, toi=toi)
        if point: out["point"] = pt
        return out

    def shapecast_all(self, casts, max_toi, capacity=0):
        """Tlas::shapecast_all: CSR (offsets, leaf, toi) sorted by (toi, instance)."""
        def fn(off, leaf, toi, cap, total):
            return lib().sq_tlas_shapecast_all(self._h, _dp(casts, "float32", 4 * self.dim), casts.shape[0], float(max_toi), _dp(off, "int64"), _dp(leaf, "int32"),
                                               _dp(toi, "float32"), cap, C.byref(total))
        return DeviceBvh._csr(self, fn, casts.shape[0], casts.device, True, capacity)

    def close(self):
        if getattr(self, "_h", None):
            lib().sq_tlas_destroy(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)


class Narrow:
    """A consumer narrow phase (GpuGeometry::narrow_wgsl, gpu/mod.rs:86-99,261):
    HIP source defining

        __device__ bool sq_narrow(const float* prim, const float* origin,
                                  const float* dir, float max_toi, float* toi);

    compiled at run time (hipRTC) into the engine's nearest-hit walk, for
    primitives of prim_floats floats in dim dimensions (the source may use
    SQ_RTC_D and SQ_RTC_PF).  Raises SqError(INVALID) with the compiler log
    when it does not compile."""

    def __init__(self, ctx: Context, dim: int, source: str, prim_floats: int):
        h = C.c_void_p()
        check(lib().sq_narrow_create(ctx.handle, dim, source.encode(), prim_floats, C.byref(h)))
        self._h, self.ctx, self.dim, self.prim_floats = h, ctx, dim, prim_floats
        _track(self)

    def build(self, boxes, prims) -> "ConsumerBvh":
        """The tree over n primitives: world boxes [n][2*dim] (host) drive the
        reference median split; prims [n][prim_floats] are what sq_narrow sees."""
        bx = _host(boxes).reshape(-1, 2 * self.dim)
        pr = _host(prims).reshape(bx.shape[0], self.prim_floats) if self.prim_floats else np.zeros((bx.shape[0], 0), np.float32)
        h = C.c_void_p()
        check(lib().sq_bvh_build_consumer(self._h, _hp(bx), _hp(pr), bx.shape[0], C.byref(h)))
        return ConsumerBvh(self, h, bx.shape[0])

    def close(self):
        if getattr(self, "_h", None):
            lib().sq_narrow_destroy(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)


class ConsumerBvh:
    """A tree over consumer primitives; answers rays through its Narrow:
    nearest hits (raycast_nearest_batch with the consumer's test_ray), all
    hits and all crossings (raycast_all / raycast_crossings)."""

    def __init__(self, narrow: Narrow, handle, n: int):
        self.narrow, self._h, self.n_prims, self.dim = narrow, handle, n, narrow.dim
        _track(self)

    def stats(self) -> dict:
        s = TreeStats()
        check(lib().sq_bvh_get_stats(self._h, C.byref(s)))
        return _stats(s)

    def raycast_nearest(self, rays, max_toi, out=None):
        import torch
        n = rays.shape[0]
        leaf = torch.empty(n, dtype=torch.int32, device=rays.device)
        toi = torch.empty(n, dtype=torch.float32, device=rays.device)
        check(lib().sq_raycast_nearest_consumer(self.narrow._h, self._h, _dp(rays, "float32", 2 * self.dim), n, float(max_toi), _dp(leaf, "int32"), _dp(toi, "float32")))
        return {"leaf": leaf, "toi": toi}

    _csr = DeviceBvh._csr

    def raycast_all(self, rays, max_toi, crossings=False, capacity=0):
        """CSR (offsets, leaf, toi), each ray's hits sorted by (toi, leaf)."""
        def fn(off, leaf, toi, cap, total):
            return lib().sq_raycast_all_consumer(self.narrow._h, self._h, _dp(rays, "float32", 2 * self.dim), rays.shape[0], float(max_toi),
                                                 int(crossings), _dp(off, "int64"), _dp(leaf, "int32"), _dp(toi, "float32"), cap, C.byref(total))
        return self._csr(fn, rays.shape[0], rays.device, True, capacity)

    def shapecast_nearest(self, casts, max_toi):
        """Nearest contact per cast [n][4D] (origin, dir, probe box min, max) through sq_narrow_cast."""
        import torch
        n = casts.shape[0]
        leaf = torch.empty(n, dtype=torch.int32, device=casts.device)
        toi = torch.empty(n, dtype=torch.float32, device=casts.device)
        check(lib().sq_shapecast_nearest_consumer(self.narrow._h, self._h, _dp(casts, "float32", 4 * self.dim), n, float(max_toi), _dp(leaf, "int32"),
                                                  _dp(toi, "float32")))
        return {"leaf": leaf, "toi": toi}

    def shapecast_all(self, casts, max_toi, capacity=0):
        def fn(off, leaf, toi, cap, total):
            return lib().sq_shapecast_all_consumer(self.narrow._h, self._h, _dp(casts, "float32", 4 * self.dim), casts.shape[0], float(max_toi),
                                                   _dp(off, "int64"), _dp(leaf, "int32"), _dp(toi, "float32"), cap, C.byref(total))
        return self._csr(fn, casts.shape[0], casts.device, True, capacity)

    def overlap(self, regions, capacity=0):
        """CSR (offsets, leaf) of ascending leaf ids through sq_narrow_overlap."""
        def fn(off, leaf, toi, cap, total):
            return lib().sq_overlap_consumer(self.narrow._h, self._h, _dp(regions, "float32", 2 * self.dim), regions.shape[0], _dp(off, "int64"), _dp(leaf, "int32"),
                                             cap, C.byref(total))
        off, leaf, _ = self._csr(fn, regions.shape[0], regions.device, False, capacity)
        return off, leaf

    def close(self):
        if getattr(self, "_h", None):
            lib().sq_bvh_destroy(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)


class Narrow64:
    """A consumer narrow phase in double (the f64 feature's QueryGeometry::
    test_ray, geometry.rs:33-43): HIP source defining

        __device__ bool sq_narrow64(const double* prim, const double* origin,
                                    const double* dir, double max_toi, double* toi);

    compiled at run time (hipRTC) into the pooled f64 walk, for primitives of
    prim_doubles doubles in dim dimensions (SQ_RTC_D and SQ_RTC_PF as for
    Narrow).  Raises SqError(INVALID) with the compiler log when it does not
    compile."""

    def __init__(self, ctx: Context, dim: int, source: str, prim_doubles: int):
        h = C.c_void_p()
        check(lib().sq_narrow64_create(ctx.handle, dim, source.encode(), prim_doubles, C.byref(h)))
        self._h, self.ctx, self.dim, self.prim_doubles = h, ctx, dim, prim_doubles
        _track(self)

    def build(self, boxes, prims) -> "ConsumerBvh64":
        """The f64 tree over n primitives: world boxes [n][2*dim] (host) drive
        the reference median split; prims [n][prim_doubles] are what
        sq_narrow64 sees."""
        bx = _host(boxes, np.float64).reshape(-1, 2 * self.dim)
        pr = (_host(prims, np.float64).reshape(bx.shape[0], self.prim_doubles) if self.prim_doubles
              else np.zeros((bx.shape[0], 0), np.float64))
        h = C.c_void_p()
        check(lib().sq_bvh64_build_consumer(self._h, _hp(bx), _hp(pr), bx.shape[0], C.byref(h)))
        return ConsumerBvh64(self, h, bx.shape[0])

    def close(self):
        if getattr(self, "_h", None):
            lib().sq_narrow64_destroy(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)


class ConsumerBvh64:
    """An f64 tree over consumer primitives; answers float64 rays through its
    Narrow64: nearest hits and all hits."""

    def __init__(self, narrow: Narrow64, handle, n: int):
        self.narrow, self._h, self.n_prims, self.dim = narrow, handle, n, narrow.dim
        _track(self)

    def stats(self) -> dict:
        s = TreeStats()
        check(lib().sq_bvh64_stats(self._h, C.byref(s)))
        return _stats(s)

    def raycast_nearest(self, rays, max_toi):
        import torch
        n = rays.shape[0]
        leaf = torch.empty(n, dtype=torch.int32, device=rays.device)
        toi = torch.empty(n, dtype=torch.float64, device=rays.device)
        check(lib().sq_raycast64_nearest_consumer(self.narrow._h, self._h, _dp(rays, "float64", 2 * self.dim), n, float(max_toi),
                                                  _dp(leaf, "int32"), _dp(toi, "float64")))
        return {"leaf": leaf, "toi": toi}

    _csr = DeviceBvh64._csr

    def raycast_all(self, rays, max_toi, capacity=0):
        """CSR (offsets, leaf, toi), each ray's hits sorted by (toi, leaf)."""
        def fn(off, leaf, toi, cap, total):
            return lib().sq_raycast64_all_consumer(self.narrow._h, self._h, _dp(rays, "float64", 2 * self.dim), rays.shape[0],
                                                   float(max_toi), _dp(off, "int64"), _dp(leaf, "int32"), _dp(toi, "float64"), cap,
                                                   C.byref(total))
        return self._csr(fn, rays.shape[0], rays.device, True, capacity)

    def raycast_crossings(self, rays, max_toi, capacity=0):
        """Bvh::raycast_crossings: CSR (offsets, leaf, toi) of every crossing
        the source's sq_narrow64_crossings reports, sorted by (toi, leaf)."""
        def fn(off, leaf, toi, cap, total):
            return lib().sq_raycast64_crossings_consumer(self.narrow._h, self._h, _dp(rays, "float64", 2 * self.dim),
                                                         rays.shape[0], float(max_toi), _dp(off, "int64"), _dp(leaf, "int32"),
                                                         _dp(toi, "float64"), cap, C.byref(total))
        return self._csr(fn, rays.shape[0], rays.device, True, capacity)

    def close(self):
        if getattr(self, "_h", None):
            lib().sq_bvh64_destroy(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)


class BatchSet:
    """Many nearest-hit batches over many scenes (all DeviceBvh, or all Tlas, of one
    dimension and kind) answered by one persistent work-stealing launch per submit().

    queries: [(scene, rays [n][2D] GPU tensor, max_toi)]; results land in
    self.leaf[i] / self.toi[i] (instance ids for Tlas scenes).  submit() only
    enqueues kernels, so it can be recorded into a CUDA/HIP graph.
    """

    def __init__(self, ctx: Context, queries):
        import torch
        if not queries:
            raise ValueError("no queries")
        tlas = isinstance(queries[0][0], Tlas)
        arr = (BatchQuery * len(queries))()
        self.leaf, self.toi, self._keep = [], [], []
        for i, (scene, rays, max_toi) in enumerate(queries):
            if isinstance(scene, Tlas) != tlas:
                raise ValueError("a batch set holds scenes of one class")
            n = rays.shape[0]
            l = torch.empty(n, dtype=torch.int32, device=rays.device)
            t = torch.empty(n, dtype=torch.float32, device=rays.device)
            self.leaf.append(l); self.toi.append(t); self._keep.append((scene, rays))
            arr[i] = BatchQuery(scene._h, _dp(rays, "float32", 2 * scene.dim).value if n else None, n, float(max_toi),
                                _dp(l).value if n else None, _dp(t).value if n else None)
        h = C.c_void_p()
        check(lib().sq_batch_create(ctx.handle, SQ_BATCH_TLAS if tlas else SQ_BATCH_BVH, C.cast(arr, C.c_void_p),
                                    len(queries), C.byref(h)))
        self._h, self.ctx = h, ctx
        _track(self)

    def submit(self):
        check(lib().sq_batch_submit(self._h))

    def valid(self) -> bool:
        """False once a scene was rebuilt into new arrays or deeper (make the set, and any graph of it, again)."""
        return lib().sq_batch_check(self._h) == 0

    def info(self) -> dict:
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().sq_batch_get_info(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return dict(chunks=a.value, grid=b.value, block=c.value)

    def close(self):
        if getattr(self, "_h", None):
            lib().sq_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)


class Proposal:
    """Learned subtree proposal for a TRIANGLE DeviceBvh or a D = 3 Tlas (no
    reference counterpart).

    An MLP predicts, per ray, the subtrees (the BVH cut at `cut_depth`; for a
    Tlas, its top tree over the instances) most likely to hold the nearest
    hit.  `raycast_nearest` walks the top-4 first to tighten t_max, then runs
    the full exact walk: results equal the scene's raycast_nearest under
    SQ_HITS_GUARDED bit for bit whatever the network predicts.  Trained on the
    GPU (`train_step`, `fit`) from labels the exact walk makes (for a Tlas, the
    hit instance ids).
    """

    def __init__(self, scene, cut_depth: int = 12, rff_scale: float = 1.0, seed: int = 0x5EED):
        h = C.c_void_p()
        self.tlas = scene if isinstance(scene, Tlas) else None
        create = lib().sq_proposal_create_tlas if self.tlas is not None else lib().sq_proposal_create
        check(create(scene.ctx.handle, scene._h, int(cut_depth), float(rff_scale), int(seed) & ((1 << 64) - 1), C.byref(h)))
        self._h, self.bvh, self.ctx, self.cut_depth = h, (None if self.tlas is not None else scene), scene.ctx, cut_depth
        _track(self)
        v = C.c_uint32()
        check(lib().sq_proposal_num_subtrees(self._h, C.byref(v)))
        self.n_subtrees = v.value
        q = C.c_uint64()
        check(lib().sq_proposal_num_params(self._h, C.byref(q)))
        self.n_params = q.value

    def shapes(self):
        """[(out, in)] of the four layers."""
        return [(256, 128), (256, 256), (256, 256), (self.n_subtrees, 256)]

    def _split(self, flat):
        out, o = [], 0
        for (m, k) in self.shapes():
            w = flat[o:o + m * k].reshape(m, k); o += m * k
            b = flat[o:o + m]; o += m
            out.append((w, b))
        return out

    def labels(self, leaf):
        import torch
        lab = torch.empty(leaf.shape[0], dtype=torch.int32, device=leaf.device)
        check(lib().sq_proposal_labels(self._h, _dp(leaf, "int32"), leaf.shape[0], _dp(lab)))
        return lab

    def train_step(self, rays, labels, lr: float = 1e-3, want_loss: bool = False):
        loss = C.c_float(0.0)
        check(lib().sq_proposal_train_step(self._h, _dp(rays, "float32", 6), _dp(labels, "int32", n=rays.shape[0]), rays.shape[0], float(lr),
                                           C.byref(loss) if want_loss else None, None))
        return loss.value if want_loss else None

    def gradients(self, rays, labels):
        """Gradients of the mean cross-entropy, [(dW, db)] per layer (fp32, no update)."""
        flat = np.empty(self.n_params, np.float32)
        check(lib().sq_proposal_train_step(self._h, _dp(rays, "float32", 6), _dp(labels, "int32", n=rays.shape[0]), rays.shape[0],
                                           0.0, None, _hp(flat)))
        return self._split(flat)

    def weights(self):
        flat = np.empty(self.n_params, np.float32)
        check(lib().sq_proposal_weights(self._h, _hp(flat), None))
        return self._split(flat)

    def set_weights(self, layers):
        flat = _host(np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for (w, b) in layers]))
        if flat.size != self.n_params:
            raise ValueError("weight count mismatch")
        check(lib().sq_proposal_weights(self._h, None, _hp(flat)))

    def features_spec(self):
        """(B float32[6][64], scene box float32[6])."""
        B = np.empty((6, 64), np.float32)
        box = np.empty(6, np.float32)
        check(lib().sq_proposal_features_spec(self._h, _hp(B), _hp(box)))
        return B, box

    def infer(self, rays, out=None):
        """Top-4 subtrees per ray, best first: int16 tensor [n][4] holding uint16 ids (0xFFFF = none)."""
        import torch
        if out is None:
            out = torch.empty((rays.shape[0], 4), dtype=torch.int16, device=rays.device)
        check(lib().sq_proposal_infer(self._h, _dp(rays, "float32", 6), rays.shape[0], _dp(out, "int16", n=rays.shape[0], per=4)))
        return out

    def raycast_nearest_into(self, rays, max_toi, top, leaf, toi, stats=None, sub=None):
        """leaf: the hit primitive (for a Tlas, the instance; sub, Tlas only, the sub-object)."""
        n = rays.shape[0]
        if self.tlas is not None:
            check(lib().sq_tlas_raycast_nearest_proposed(self._h, _dp(rays, "float32", 6), n, float(max_toi),
                                                         _dp(top, "int16", n=n, per=4), _dp(leaf, "int32", n=n),
                                                         _dp(toi, "float32", n=n), _dp(sub, "int32", n=n),
                                                         _dp(stats, "int64", n=3)))
            return
        if sub is not None:
            raise ValueError("sub-objects are reported for a Tlas proposal only")
        check(lib().sq_raycast_nearest_proposed(self._h, _dp(rays, "float32", 6), n, float(max_toi),
                                                _dp(top, "int16", n=n, per=4), _dp(leaf, "int32", n=n),
                                                _dp(toi, "float32", n=n), _dp(stats, "int64", n=3)))

    def raycast_nearest(self, rays, max_toi, top=None, count=False):
        """Exact nearest hit walking `top` (None = the plain walk) first: dict(leaf, toi[, sub][, seeded, nodes,
        prims]); for a Tlas, leaf is the instance and sub the sub-object, and only `seeded` is counted."""
        import torch
        n = rays.shape[0]
        leaf = torch.empty(n, dtype=torch.int32, device=rays.device)
        toi = torch.empty(n, dtype=torch.float32, device=rays.device)
        sub = torch.empty(n, dtype=torch.int32, device=rays.device) if self.tlas is not None else None
        st = torch.zeros(3, dtype=torch.int64, device=rays.device) if count else None
        self.raycast_nearest_into(rays, max_toi, top, leaf, toi, st, sub)
        out = dict(leaf=leaf, toi=toi)
        if sub is not None:
            out["sub"] = sub
        if count:
            s = st.cpu().tolist()
            out.update(seeded=s[0], nodes=s[1], prims=s[2])
        return out

    def fit(self, rays, labels, epochs: int = 1, batch: int = 65536, lr: float = 1e-3, seed: int = 1, log=None):
        """Adam over shuffled mini-batches of (rays, labels) (GPU tensors); returns the mean loss per epoch."""
        import torch
        g = torch.Generator(device=rays.device)
        g.manual_seed(seed)
        n = rays.shape[0]
        losses = []
        for e in range(epochs):
            perm = torch.randperm(n, device=rays.device, generator=g)
            tot, cnt = 0.0, 0
            for s0 in range(0, n - batch + 1, batch):
                idx = perm[s0:s0 + batch]
                r = rays.index_select(0, idx).contiguous()
                l = labels.index_select(0, idx).contiguous()
                last = s0 + 2 * batch > n
                v = self.train_step(r, l, lr, want_loss=(s0 // batch) % 16 == 0 or last)
                if v is not None:
                    tot += v; cnt += 1
            losses.append(tot / max(cnt, 1))
            if log:
                log(f"epoch {e}: loss {losses[-1]:.4f}")
        return losses

    def close(self):
        if getattr(self, "_h", None):
            lib().sq_proposal_destroy(self._h)
            self._h = None

    def __del__(self):
        _finalize(self)


def mfma_gemm(ctx: Context, A, B, splits: int = 1):
    """C = bf16(A) @ bf16(B)^T in fp32 on the matrix cores (the training GEMM); A [M][K], B [N][K] GPU fp32."""
    import torch
    M, K = A.shape
    N = B.shape[0]
    Cm = torch.empty((M, N), dtype=torch.float32, device=A.device)
    check(lib().sq_debug_gemm(ctx.handle, _dp(A), _dp(B), M, N, K, splits, _dp(Cm)))
    return Cm
