// pool.h -- the pooled walk shared by the single-scene kernels (engine.hip)
// and the persistent batch kernel (batch.hip).  Device code only.
#pragma once

#include <type_traits>

#include "engine_internal.h"

// Pause a wave's descents once fewer than `below` lanes still descend while
// other lanes of the wave hold queries (at a leaf, or finished and waiting for a
// refill): those lanes then get on with their leaf tests and new queries, and
// the long descents resume in the next step.
struct WavePause {
  uint32_t below, active;
  __device__ __forceinline__ bool operator()() const {
    const uint32_t n = (uint32_t)__popcll(__ballot(1));
    return n < below && n < active;
  }
};

// Next value of a wave's atomic counter (called with every lane active).
__device__ __forceinline__ uint32_t wave_fetch_add(uint32_t* counter) {
  uint32_t c = 0;
  if ((threadIdx.x & 63u) == 0) c = atomicAdd(counter, 1u);
  return __builtin_amdgcn_readfirstlane(c);
}

// Query ranges a wave walks.  OnePool: one fixed range.  Chunks: ranges of
// `chunk` queries taken from an atomic counter until n is covered (work
// stealing across the whole grid).
struct OnePool {
  uint64_t first, end;
  bool given = false;
  __device__ __forceinline__ bool get(uint64_t& next, uint64_t& e) {
    if (given) return false;
    given = true;
    next = first;
    e = end;
    return first < end;
  }
};
struct Chunks {
  uint32_t* counter;
  uint64_t n, chunk;
  uint64_t home;          // this wave's first chunk; later ones come from the counter, after all home chunks
  uint64_t waves;
  bool started = false;
  __device__ __forceinline__ bool get(uint64_t& next, uint64_t& e) {
    uint64_t c;
    if (!started) { started = true; c = home; }
    else c = waves + (uint64_t)wave_fetch_add(counter);
    const uint64_t f = c * chunk;
    if (f >= n) return false;
    next = f;
    e = f + chunk < n ? f + chunk : n;
    return true;
  }
};

// One wave walks the query ranges src hands it (wave-uniform).  Lanes take
// queries from the current range in order; a lane whose query finishes takes
// the next one (ballot + popcount), and the wave takes a new range when this
// one runs out, so lanes stay busy while long queries run.  Each step walks to one leaf and tests it (walk_step), or stops early
// when fewer than pause_below lanes are still descending (WavePause).  The stack is
// the caller's LDS ring spilling to a global column.  With `order`, pool
// position i is query order[i] (rays read and results written there).
template <int D, int KIND, bool ANY, bool FILT, bool CAST, class Src>
__device__ __forceinline__ void pool_walk(const SceneView& s, const float* __restrict__ q, Src src, float max_toi,
                                          const float* __restrict__ tmax, const Filter& f, const NearestOut& out,
                                          const RingStack& st, const uint32_t* __restrict__ order, uint32_t pause_below) {
  const uint32_t lane = threadIdx.x & 63u;
  using Pol = typename std::conditional<CAST, CastPolicy<D, KIND>, RayPolicy<D, KIND>>::type;
  constexpr int QS = CAST ? 4 * D : 2 * D;
  constexpr int PS = Prim<D, KIND>::STRIDE;
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  Pol pol;
  WalkState w;
  float lim0 = 0.0f;
  uint32_t rid = NO_LEAF;
  uint64_t next = 0, end = 0;
  bool exhausted = !src.get(next, end);
  while (true) {
    bool need = rid == NO_LEAF && !exhausted;
    uint64_t m = __ballot(need);
    while (m) {
      // hand the pool's next queries to the lanes that need one, in lane
      // order; take a new range when this one runs out
      if (next >= end) {
        if (!src.get(next, end)) { exhausted = true; break; }
        continue;
      }
      const uint64_t avail = end - next;
      const uint64_t rank = (uint64_t)__popcll(m & lt_mask);
      if (need && rank < avail) {
        need = false;
        {
          const uint64_t my = next + rank;
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
        }
      }
      const uint64_t taken = (uint64_t)__popcll(m);
      next += taken < avail ? taken : avail;
      m = __ballot(need);
    }
    if (__ballot(!(exhausted && rid == NO_LEAF)) == 0) break;
    const uint32_t active = (uint32_t)__popcll(__ballot(rid != NO_LEAF));
    if (rid == NO_LEAF) continue;
    if (!walk_step<D, PS, ANY, FILT>(s, pol, f, st, w, nullptr, nullptr, WavePause{pause_below, active})) continue;
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
