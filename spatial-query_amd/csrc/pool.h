// pool.h -- the pooled walk shared by the single-scene kernels (engine.hip)
// and the persistent batch kernel (batch.hip).  Device code only.
#pragma once

#include <type_traits>

#include "engine_internal.h"

// One wave walks the queries [next, end) (wave-uniform bounds).  Lanes take
// queries from this pool in order; a lane whose query finishes takes the next
// one (ballot + popcount, no atomics), so lanes stay busy while long queries
// run.  Each step descends to one leaf and tests it (walk_step).  The stack is
// the caller's LDS ring spilling to a global column.  With `order`, pool
// position i is query order[i] (rays read and results written there).
template <int D, int KIND, bool ANY, bool FILT, bool CAST>
__device__ __forceinline__ void pool_walk(const SceneView& s, const float* __restrict__ q, uint64_t next, const uint64_t end,
                                          float max_toi, const float* __restrict__ tmax, const Filter& f, const NearestOut& out,
                                          const RingStack& st, const uint32_t* __restrict__ order) {
  const uint32_t lane = threadIdx.x & 63u;
  using Pol = typename std::conditional<CAST, CastPolicy<D, KIND>, RayPolicy<D, KIND>>::type;
  constexpr int QS = CAST ? 4 * D : 2 * D;
  constexpr int PS = Prim<D, KIND>::STRIDE;
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  Pol pol;
  WalkState w;
  float lim0 = 0.0f;
  uint32_t rid = NO_LEAF;
  bool exhausted = next >= end;
  while (true) {
    const bool need = rid == NO_LEAF && !exhausted;
    const uint64_t m = __ballot(need);
    if (m) {
      if (need) {
        const uint64_t my = next + __popcll(m & lt_mask);
        if (my < end) {
          rid = order ? order[my] : (uint32_t)my;
          pol.setup(q + (size_t)rid * QS);
          lim0 = tmax ? tmax[rid] : max_toi;
          if (!walk_begin<D>(s, pol, lim0, w)) {
            if (ANY) out.any[rid] = 0;
            else {
              out.leaf[rid] = -1; out.toi[rid] = SQ_INF;
              if (out.point) for (int k = 0; k < D; ++k) out.point[(size_t)rid * D + k] = 0.0f;
              if (!CAST && out.normal) for (int k = 0; k < D; ++k) out.normal[(size_t)rid * D + k] = 0.0f;
              if (out.sub) out.sub[rid] = -1;
            }
            rid = NO_LEAF;
          }
        } else {
          exhausted = true;
        }
      }
      next += __popcll(m);
    }
    if (__ballot(!(exhausted && rid == NO_LEAF)) == 0) break;
    if (rid == NO_LEAF) continue;
    if (!walk_step<D, PS, ANY, FILT>(s, pol, f, st, w)) continue;
    // finished: write the result of query rid
    const NearestResult& h = w.res;
    const bool hit = h.leaf != NO_LEAF;
    if (ANY) {
      out.any[rid] = hit ? 1 : 0;
    } else {
      out.leaf[rid] = hit ? (int32_t)h.leaf : -1;
      out.toi[rid] = h.t;
      if (out.point) {
#pragma unroll
        for (int k = 0; k < D; ++k) out.point[(size_t)rid * D + k] = hit ? pol.r.o[k] + pol.r.d[k] * h.t : 0.0f;
      }
      if (!CAST && out.normal) {
        float nv[D];
        if (hit) hit_normal<D, KIND>(pol.r, s.prims + (size_t)h.pos * PS, h.t, lim0, nv);
        else for (int k = 0; k < D; ++k) nv[k] = 0.0f;
#pragma unroll
        for (int k = 0; k < D; ++k) out.normal[(size_t)rid * D + k] = nv[k];
      }
      if (out.sub) out.sub[rid] = (hit && KIND != KIND_AABB) ? (int32_t)h.leaf : -1;
    }
    rid = NO_LEAF;
  }
}
