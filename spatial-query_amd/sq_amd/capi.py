"""ctypes declarations of include/sq.h."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(os.path.dirname(_HERE), "lib", "libsq_amd.so")

SQ_OK, SQ_ERR_INVALID, SQ_ERR_NO_DEVICE, SQ_ERR_HIP, SQ_ERR_UNSUPPORTED, SQ_ERR_CAPACITY = range(6)
SQ_PRIM_AABB, SQ_PRIM_SPHERE, SQ_PRIM_TRIANGLE = 0, 1, 2
SQ_BATCH_BVH, SQ_BATCH_TLAS = 0, 1
SQ_HITS_REFERENCE, SQ_HITS_GUARDED = 0, 1


class SqError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sq error {code}: {msg}")
        self.code = code


class Filter(C.Structure):
    _fields_ = [("layer_mask", C.c_uint64), ("exclude", C.c_int64)]


class TreeStats(C.Structure):
    _fields_ = [("node_count", C.c_uint64), ("leaf_count", C.c_uint64), ("prim_count", C.c_uint64),
                ("max_depth", C.c_uint64), ("bytes", C.c_uint64), ("quality", C.c_float)]


class BatchQuery(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("rays", C.c_void_p), ("n", C.c_uint64), ("max_toi", C.c_float),
                ("out_leaf", C.c_void_p), ("out_toi", C.c_void_p)]


class Tuning(C.Structure):
    _fields_ = [("block", C.c_int), ("rays_per_lane", C.c_int), ("lds_stack", C.c_int), ("ray_order", C.c_int),
                ("descend_min", C.c_int), ("split_depth", C.c_int), ("wide", C.c_int), ("packet", C.c_int),
                ("lookahead", C.c_int), ("lookahead_pause", C.c_int)]


P, U64, I64, F32, I32, U32 = C.c_void_p, C.c_uint64, C.c_int64, C.c_float, C.c_int, C.c_uint32
F64 = C.c_double
PP = C.POINTER(C.c_void_p)

# name -> argtypes (all return int unless listed in _RESTYPE)
SIGNATURES = {
    "sq_last_error": [],
    "sq_version": [],
    "sq_ctx_create": [I32, PP],
    "sq_ctx_destroy": [P],
    "sq_ctx_synchronize": [P],
    "sq_ctx_stream": [P],
    "sq_ctx_set_stream": [P, P],
    "sq_ctx_reset_stream": [P],
    "sq_tlas_build": [P, I32, I32, U64, P, P, P, P, U32, P, P, PP],
    "sq_tlas_destroy": [P],
    "sq_tlas_refit": [P, P, P, P],
    "sq_tlas_needs_rebuild": [P, F32, C.POINTER(C.c_int)],
    "sq_tlas_get_stats": [P, C.POINTER(TreeStats)],
    "sq_tlas_download_bounds": [P, P],
    "sq_tlas_raycast_nearest": [P, P, U64, F32, P, P, P, P, P, P],
    "sq_tlas_raycast_any": [P, P, U64, F32, P, P],
    "sq_tlas_raycast_all": [P, P, U64, F32, P, P, P, U64, C.POINTER(C.c_uint64)],
    "sq_tlas_overlap": [P, P, U64, P, P, U64, C.POINTER(C.c_uint64)],
    "sq_tlas_shapecast_nearest": [P, P, U64, F32, P, P, P],
    "sq_tlas_shapecast_all": [P, P, U64, F32, P, P, P, U64, C.POINTER(C.c_uint64)],
    "sq_ctx_set_block": [P, I32],
    "sq_batch_create": [P, I32, P, U32, PP],
    "sq_batch_submit": [P],
    "sq_batch_check": [P],
    "sq_batch_get_info": [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)],
    "sq_batch_destroy": [P],
    "sq_proposal_create": [P, P, I32, F32, U64, PP],
    "sq_proposal_create_tlas": [P, P, I32, F32, U64, PP],
    "sq_proposal_destroy": [P],
    "sq_proposal_num_subtrees": [P, C.POINTER(C.c_uint32)],
    "sq_proposal_num_params": [P, C.POINTER(C.c_uint64)],
    "sq_proposal_labels": [P, P, U64, P],
    "sq_proposal_train_step": [P, P, P, U64, F32, C.POINTER(C.c_float), P],
    "sq_proposal_weights": [P, P, P],
    "sq_proposal_features_spec": [P, P, P],
    "sq_proposal_infer": [P, P, U64, P],
    "sq_raycast_nearest_proposed": [P, P, U64, F32, P, P, P, P],
    "sq_tlas_raycast_nearest_proposed": [P, P, U64, F32, P, P, P, P, P],
    "sq_debug_gemm": [P, P, P, I32, I32, I32, I32, P],
    "sq_ctx_set_tuning": [P, C.POINTER(Tuning)],
    "sq_ctx_set_hit_mode": [P, I32],
    "sq_ctx_get_hit_mode": [P],
    "sq_ctx_get_tuning": [P, C.POINTER(Tuning)],
    "sq_ctx_last_kernel_ms": [P, C.POINTER(C.c_float)],
    "sq_tree_build": [I32, I32, P, U64, PP],
    "sq_tree_from_nodes": [I32, P, P, P, P, U64, P, U64, PP],
    "sq_tree_refit": [P, I32, P],
    "sq_tree_get_stats": [P, C.POINTER(TreeStats)],
    "sq_tree_export": [P, P, P, P, P, P],
    "sq_tree_destroy": [P],
    "sq_bvh_upload": [P, P, I32, P, PP],
    "sq_bvh_build": [P, I32, I32, P, U64, PP],
    "sq_bvh_destroy": [P],
    "sq_bvh_set_layers": [P, P],
    "sq_bvh_get_stats": [P, C.POINTER(TreeStats)],
    "sq_bvh_download_bounds": [P, P],
    "sq_bvh_export": [P, P, P, P, P, P],
    "sq_bvh_refit": [P, P],
    "sq_bvh_build_device": [P, I32, I32, P, U64, PP],
    "sq_bvh_rebuild": [P, P],
    "sq_bvh_surface_sum": [P, C.POINTER(C.c_double)],
    "sq_bvh_refit_auto": [P, P, F32, C.POINTER(C.c_int)],
    "sq_narrow_create": [P, I32, C.c_char_p, U32, PP],
    "sq_narrow_destroy": [P],
    "sq_bvh_build_consumer": [P, P, P, U64, PP],
    "sq_raycast_nearest_consumer": [P, P, P, U64, F32, P, P],
    "sq_raycast_all_consumer": [P, P, P, U64, F32, I32, P, P, P, U64, C.POINTER(C.c_uint64)],
    "sq_shapecast_nearest_consumer": [P, P, P, U64, F32, P, P],
    "sq_shapecast_all_consumer": [P, P, P, U64, F32, P, P, P, U64, C.POINTER(C.c_uint64)],
    "sq_overlap_consumer": [P, P, P, U64, P, P, U64, C.POINTER(C.c_uint64)],
    "sq_raycast_nearest": [P, P, U64, F32, P, C.POINTER(Filter), P, P, P, P, P],
    "sq_raycast_nearest_noalloc": [P, P, U64, F32, P, C.POINTER(Filter), P, P, P, P, P],
    "sq_raycast_any": [P, P, U64, F32, P, C.POINTER(Filter), P],
    "sq_raycast_nearest_stats": [P, P, U64, F32, P, C.POINTER(Filter), P, P, P],
    "sq_raycast_all_stats": [P, P, U64, F32, C.POINTER(Filter), P],
    "sq_raycast_all": [P, P, U64, F32, C.POINTER(Filter), I32, P, P, P, U64, C.POINTER(C.c_uint64)],
    "sq_overlap": [P, P, U64, C.POINTER(Filter), P, P, U64, C.POINTER(C.c_uint64)],
    "sq_shapecast_nearest": [P, P, U64, F32, P, C.POINTER(Filter), P, P, P],
    "sq_shapecast_all": [P, P, U64, F32, P, C.POINTER(Filter), P, P, P, U64, C.POINTER(C.c_uint64)],
    "sq_required_work": [P, P, U64, P, P, P],
    "sq_bvh64_create": [P, I32, I32, P, U64, PP],
    "sq_bvh64_destroy": [P],
    "sq_bvh64_stats": [P, C.POINTER(TreeStats)],
    "sq_bvh64_set_layers": [P, P],
    "sq_bvh64_raycast_nearest": [P, P, U64, F64, P, C.POINTER(Filter), P, P, P, P],
    "sq_bvh64_raycast_any": [P, P, U64, F64, P, C.POINTER(Filter), P],
    "sq_bvh64_raycast_all": [P, P, U64, F64, C.POINTER(Filter), P, P, P, U64, C.POINTER(C.c_uint64)],
    "sq_bvh64_overlap": [P, P, U64, C.POINTER(Filter), P, P, U64, C.POINTER(C.c_uint64)],
    "sq_narrow64_create": [P, I32, C.c_char_p, U32, PP],
    "sq_narrow64_destroy": [P],
    "sq_bvh64_build_consumer": [P, P, P, U64, PP],
    "sq_raycast64_nearest_consumer": [P, P, P, U64, F64, P, P],
    "sq_raycast64_all_consumer": [P, P, P, U64, F64, P, P, P, U64, C.POINTER(C.c_uint64)],
    "sq_raycast64_crossings_consumer": [P, P, P, U64, F64, P, P, P, U64, C.POINTER(C.c_uint64)],
    "sq_tlas64_create": [P, I32, U64, P, P, P, PP],
    "sq_tlas64_create_kind": [P, I32, I32, U64, P, P, P, PP],
    "sq_tlas64_raycast_all": [P, P, U64, F64, P, P, P, U64, P],
    "sq_tlas64_overlap": [P, P, U64, P, P, U64, P],
    "sq_tlas64_shapecast_nearest": [P, P, U64, F64, P, P, P],
    "sq_tlas64_shapecast_all": [P, P, U64, F64, P, P, P, U64, P],
    "sq_tlas64_destroy": [P],
    "sq_tlas64_refit": [P, P, P, P],
    "sq_tlas64_needs_rebuild": [P, C.c_double, C.POINTER(C.c_int)],
    "sq_bvh64_refit": [P, P],
    "sq_tlas64_stats": [P, C.POINTER(TreeStats)],
    "sq_tlas64_raycast_nearest": [P, P, U64, F64, P, P, P, P],
    "sq_tlas64_raycast_any": [P, P, U64, F64, P, P],
    "sq_bvh64_shapecast_nearest": [P, P, U64, F64, P, C.POINTER(Filter), P, P, P],
    "sq_bvh64_shapecast_all": [P, P, U64, F64, P, C.POINTER(Filter), P, P, P, U64, C.POINTER(C.c_uint64)],
}
_RESTYPE = {"sq_last_error": C.c_char_p, "sq_ctx_stream": C.c_void_p}

_lib = None


def library_path() -> str:
    return _LIB


def lib():
    """The loaded engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise ImportError(f"{_LIB} is missing: build it with `make -C spatial-query_amd` "
                              "(or __graft_entry__.build()); the engine has no CPU fallback")
        L = C.CDLL(_LIB)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, C.c_int)
        _lib = L
    return _lib


def check(rc: int):
    if rc != SQ_OK:
        raise SqError(rc, lib().sq_last_error().decode(errors="replace"))
