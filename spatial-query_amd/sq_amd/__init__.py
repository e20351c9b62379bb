"""Python binding of the MI355X batched BVH query engine (include/sq.h).

The names mirror the reference surface (crates/spatial-query/src/lib.rs:68-73):
`Tree.build` is `Bvh::build` on the host (the canonical tree), `DeviceBvh` is
the device backend that the reference's `GpuRaycaster` / `QueryBackend` slot
holds, `QueryFilter` is the reference filter.  Query inputs and outputs are
torch tensors on the GPU; results are bit-identical to the reference CPU path
(the default hit mode, SQ_HITS_REFERENCE).

The package loads `lib/libsq_amd.so` (built by `make -C spatial-query_amd`) and
raises when it is missing: there is no CPU fallback.
"""
from .capi import (SQ_HITS_GUARDED, SQ_HITS_REFERENCE, SQ_PRIM_AABB, SQ_PRIM_SPHERE, SQ_PRIM_TRIANGLE, SqError, lib,
                   library_path)
from .api import (BatchSet, ConsumerBvh, ConsumerBvh64, Context, DeviceBvh, DeviceBvh64, Narrow, Narrow64, Proposal, QueryFilter, Tlas, Tlas64, Tree, mfma_gemm, prim_stride)

AABB, SPHERE, TRIANGLE = SQ_PRIM_AABB, SQ_PRIM_SPHERE, SQ_PRIM_TRIANGLE

__all__ = ["BatchSet", "ConsumerBvh", "ConsumerBvh64", "Context", "DeviceBvh", "DeviceBvh64", "Narrow", "Narrow64", "Proposal", "mfma_gemm", "Tlas", "Tlas64", "QueryFilter", "Tree", "SqError", "lib", "library_path", "prim_stride",
           "AABB", "SPHERE", "TRIANGLE", "SQ_HITS_REFERENCE", "SQ_HITS_GUARDED"]
