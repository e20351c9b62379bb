// pool.h -- the pooled walk shared by the single-scene kernels (engine.hip),
// the persistent batch kernel (batch.hip) and consumer narrow-phase kernels
// compiled at run time (rtc.hip).  Device code only; needs nothing beyond
// traverse.h, so that hipRTC can compile it too.
#pragma once

#include <type_traits>

#include "traverse.h"

namespace sqd {

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

// A persistent launch's pool counters {next chunk, retired waves} start at
// zero.  Every wave of the grid calls pool_retire once, after its last fetch
// (the one that found the pool empty, whose value it has already used); the
// wave that retires last finds every fetch of the launch done and sets both
// words back to zero.  The next launch in stream order therefore finds them
// cleared without a clearing launch of its own.
__device__ __forceinline__ void pool_retire(uint32_t* counter) {
  if ((threadIdx.x & 63u) == 0) {
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    if (atomicAdd(counter + 1, 1u) == waves - 1u) {
      atomicExch(counter, 0u);
      atomicExch(counter + 1, 0u);
    }
  }
}

// Query ranges a wave walks.  OnePool: one fixed range.  Chunks: ranges of
// `chunk` queries taken from an atomic counter until n is covered (work
// stealing across the whole grid).  ItemChunks: the chunks of one batch of a
// batch set (numbered chunk0 .. chunk_end - 1 in the set's shared counter),
// starting from a chunk the wave already holds; a fetched chunk of a later
// batch ends this batch's walk and is left in *carry for the next one.
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
struct ItemChunks {
  uint32_t* counter;
  uint32_t chunk0, chunk_end;
  uint64_t n, chunk;       // the batch's queries, queries per chunk
  uint32_t* carry;         // in: the chunk held; out: the first chunk fetched past this batch
  bool started = false;
  __device__ __forceinline__ bool get(uint64_t& next, uint64_t& e) {
    uint32_t c = *carry;
    if (started) {
      c = wave_fetch_add(counter);
      *carry = c;
    }
    started = true;
    if (c < chunk0 || c >= chunk_end) return false;
    const uint64_t f = (uint64_t)(c - chunk0) * chunk;
    next = f;
    e = f + chunk < n ? f + chunk : n;
    return true;
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
// when fewer than pause_below lanes are still descending (WavePause), so
// that finished lanes can take new queries and lanes waiting at a leaf test
// it instead of waiting for the longest descent.  The stack is
// the caller's LDS ring spilling to a global column.  With `order`, pool
// position i is query order[i] (rays read and results written there).
// The walk for any query policy Pol (setup / root / pair / narrow over
// primitive records of PS floats, queries of QS floats).  KIND picks what the
// optional point / normal / sub outputs mean (the built-in kinds); CAST marks
// shape casts (no normal); REF walks the reference's order (walk_step).
template <int D, int KIND, bool ANY, bool FILT, bool CAST, class Pol, int QS, int PS, class Src, bool QUAD = false,
          bool REF = false, bool SPEC = false>
__device__ __forceinline__ void pool_walk_with(const SceneView& s, const float* __restrict__ q, Src src, float max_toi,
                                               const float* __restrict__ tmax, const Filter& f, const NearestOut& out,
                                               const RingStack& st, const uint32_t* __restrict__ order, uint32_t pause_below) {
  static_assert(!SPEC || (REF && !ANY), "the look-ahead walk is the reference-order nearest walk");
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  Pol pol;
  typename std::conditional<SPEC, SpecState, WalkState>::type w;
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
          bool live;
          if constexpr (SPEC) live = spec_begin<D>(s, pol, lim0, w);
          else live = walk_begin<D>(s, pol, lim0, w);
          if (!live) {
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
    bool done;
    if constexpr (SPEC) done = walk_step_spec<D, PS, FILT, QUAD>(s, pol, f, st, w, WavePause{pause_below, active});
    else done = walk_step<D, PS, ANY, FILT, false, false, QUAD, REF>(s, pol, f, st, w, nullptr, nullptr, WavePause{pause_below, active});
    if (!done) continue;
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

// The built-in kinds: rays (RayPolicy) or shape casts (CastPolicy); QUAD
// walks the 4-wide records (s.quads).  EXACT: the reference's answer for every
// query (unguarded narrow phases; nearest queries whose hits may leave their
// box walk the reference's order); else the guarded kinds near-first.  SPEC:
// the reference-order walks take the look-ahead form (walk_step_spec).
template <int D, int KIND, bool ANY, bool FILT, bool CAST, class Src, bool QUAD = false, bool EXACT = true, bool SPEC = false>
__device__ __forceinline__ void pool_walk(const SceneView& s, const float* __restrict__ q, Src src, float max_toi,
                                          const float* __restrict__ tmax, const Filter& f, const NearestOut& out,
                                          const RingStack& st, const uint32_t* __restrict__ order, uint32_t pause_below) {
  using Pol = typename std::conditional<CAST, CastPolicy<D, KIND, float, !EXACT>, RayPolicy<D, KIND, !EXACT>>::type;
  constexpr bool REF = EXACT && !ANY && !Pol::INSIDE;
  pool_walk_with<D, KIND, ANY, FILT, CAST, Pol, CAST ? 4 * D : 2 * D, Prim<D, KIND>::STRIDE, Src, QUAD, REF, SPEC && REF>(
      s, q, src, max_toi, tmax, f, out, st, order, pause_below);
}

// ---------------------------------------------------------------------------
// Two-level walk, resumable (Tlas::raycast_nearest, tlas.rs:104-154, the
// order of walk_tlas in traverse.h made stepwise): a step descends the top
// tree to a leaf, or tests the next instances of that leaf (box instances
// directly; a mesh instance starts a walk of its BLAS with the local ray), or
// takes one walk_step in the current BLAS.  A finished BLAS walk's hit goes
// through the (toi, instance) order (with GUARD after instance_test's guard);
// the limit the BLAS walk prunes with is the top-level best when it started,
// as in instance_test.  EXACT walks the top level in the reference's order
// (order_top) and each BLAS as the reference's Bvh::raycast_nearest; else both
// guarded, near-first.  Both stacks are LDS rings spilling to device memory.
// ---------------------------------------------------------------------------

template <int D, bool EXACT = true> struct TlasWalk {
  uint32_t cur, sp;            // top level: next ref, stack depth
  float limit;
  TlasHit res;
  uint32_t first, cnt, j;      // the top leaf range under test, j its next instance
  uint32_t phase;              // 0 descend the top, 1 test the range, 2 walk a BLAS, 3 pop the top
  uint32_t inst, mesh;         // the mesh instance being walked
  RayPolicy<D, KIND_TRI, !EXACT> pol;  // its local ray
  WalkState wb;
};

template <int D, bool EXACT>
__device__ __forceinline__ bool tlas_begin(const TlasView& tv, const RayD<D>& r, float max_toi, TlasWalk<D, EXACT>& w) {
  w.res = TlasHit{SQ_INF, NO_LEAF, NO_LEAF, NO_LEAF};
  w.limit = max_toi;
  w.sp = 0;
  w.phase = 0;
  w.cur = fbits(tv.top.root[2 * D]);
  float te, tx;
  int ax;
  return slab_exact<D>(r, tv.top.root, tv.top.root + D, max_toi, te, tx, ax);
}

template <int D, bool EXACT>
__device__ __forceinline__ void tlas_take(TlasWalk<D, EXACT>& w, float t, uint32_t inst, uint32_t pos, uint32_t sub) {
  if (w.res.inst == NO_LEAF || hit_before(t, inst, w.res.t, w.res.inst)) {
    w.res = TlasHit{t, inst, pos, sub};
    w.limit = t;
  }
}

// One step of ray r; true when the ray is finished (result in w.res).  ANY
// finishes at the first accepted instance hit (walk_tlas's ANY).  SKIP treats
// the top-tree child refs skip[0..3] as missed (subtrees already walked).
template <int D, int IK, bool ANY, bool EXACT, bool SKIP = false, class P>
__device__ __forceinline__ bool tlas_step(const TlasView& tv, const RayD<D>& r, const RingStack& st, const RingStack& sb,
                                          TlasWalk<D, EXACT>& w, const P& pause, const uint32_t* skip = nullptr) {
  constexpr int NF = (D + 1) * 4;
  const SceneView& s = tv.top;
  if constexpr (IK == INST_MESH) {
    if (w.phase == 2) {
    const Filter nof{nullptr, ~0ull, -1};
    if (!walk_step<D, Prim<3, KIND_TRI>::STRIDE, false, false, false, false, false, EXACT>(tv.blas[w.mesh], w.pol, nof, sb,
                                                                                          w.wb, nullptr, nullptr, pause))
      return false;
    if (w.wb.res.leaf != NO_LEAF) {
      const float t = w.wb.res.t;
      bool keep = true;
      if constexpr (!EXACT) {
        // instance_test's guard: toi inside the instance's world slab at the limit the BLAS walk used
        const float* wx = tv.world + (size_t)w.inst * 2 * D;
        float te, tx;
        int a;
        keep = slab_exact<D>(r, wx, wx + D, w.limit, te, tx, a) && te <= t && t <= tx;
      }
      if (keep) {
        tlas_take<D>(w, t, w.inst, w.wb.res.pos, w.wb.res.leaf);
        if (ANY) return true;
      }
    }
    ++w.j;
    w.phase = 1;
    }
  }
  if (w.phase == 1) {
    while (w.j < w.cnt) {
      const uint32_t id = s.ids[w.first + w.j];
      if constexpr (IK != INST_MESH) {
        TlasHit h;
        if (instance_test<D, IK, !EXACT>(tv, r, id, w.limit, Stack{nullptr, nullptr, 0}, h)) {
          tlas_take<D>(w, h.t, h.inst, h.pos, h.sub);
          if (ANY) return true;
        }
        ++w.j;
      } else {
        const float* x = tv.xf + (size_t)id * tv.xs;
        float ol[D], dl[D], src[2 * D];
        iso_inverse<D>(x, x + D * D, r, ol, dl);
        const float inv_s = x[D * D + D];
#pragma unroll
        for (int k = 0; k < D; ++k) { src[k] = ol[k] * inv_s; src[D + k] = dl[k] * inv_s; }
        w.pol.setup(src);
        w.mesh = fbits(x[D * D + D + 1]);
        if (!walk_begin<D>(tv.blas[w.mesh], w.pol, w.limit, w.wb)) { ++w.j; continue; }
        w.inst = id;
        w.phase = 2;
        return false;                 // the BLAS walk starts with the next step
      }
    }
    w.phase = 3;
  }
  if (w.phase == 3) {
    bool got = false;
    while (w.sp > 0) {
      uint32_t ref;
      float tt;
      st.pop(w.sp, ref, tt);
      if (tt <= w.limit) { w.cur = ref; got = true; break; }
    }
    if (!got) return true;
    w.phase = 0;
  }
  // phase 0: descend the top tree to a leaf
  while (!(w.cur & LEAF_BIT)) {
    float nd[NF];
    fetch16<D + 1>(s.nodes + (size_t)w.cur * NF, nd);
    bool ha, hb, a_first;
    float ta, tb;
    order_top<D, EXACT, ANY>(r, nd, w.limit, ha, hb, ta, tb, a_first);
    const uint32_t ra = fbits(nd[4 * D + 0]), rb = fbits(nd[4 * D + 1]);
    if constexpr (SKIP) {
#pragma unroll
      for (int k = 0; k < 4; ++k) { ha = ha && ra != skip[k]; hb = hb && rb != skip[k]; }
    }
    if (ha && hb) {
      st.push(w.sp, a_first ? rb : ra, a_first ? tb : ta);
      w.cur = a_first ? ra : rb;
    } else if (ha) {
      w.cur = ra;
    } else if (hb) {
      w.cur = rb;
    } else {
      bool got = false;
      while (w.sp > 0) {
        uint32_t ref;
        float tt;
        st.pop(w.sp, ref, tt);
        if (tt <= w.limit) { w.cur = ref; got = true; break; }
      }
      if (!got) return true;
    }
    if (pause()) return false;
  }
  w.first = w.cur & FIRST_MASK;
  w.cnt = ((w.cur >> 27) & 15u) + 1u;
  w.j = 0;
  w.phase = 1;
  return false;
}

// One wave walks the ray ranges src hands it through the two-level scene
// (lanes refilled as rays finish, as pool_walk).  Outputs as k_tlas_nearest:
// any, or leaf (instance) and toi with optional point, normal, sub.  EXACT as
// tlas_step.
template <int D, int IK, bool ANY, class Src, bool EXACT = true>
__device__ __forceinline__ void pool_walk_tlas(const TlasView& tv, const float* __restrict__ q, Src src, float max_toi,
                                               const float* __restrict__ tmax, const NearestOut& out, const RingStack& st,
                                               const RingStack& sb, uint32_t pause_below) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  RayD<D> r;
  TlasWalk<D, EXACT> w;
  float lim0 = 0.0f;
  uint32_t rid = NO_LEAF;
  uint64_t next = 0, end = 0;
  bool exhausted = !src.get(next, end);
  auto write = [&](uint32_t i) {
    const bool hit = w.res.inst != NO_LEAF;
    if (ANY) {
      out.any[i] = hit ? 1 : 0;
      return;
    }
    out.leaf[i] = hit ? (int32_t)w.res.inst : -1;
    out.toi[i] = w.res.t;
    if (out.point) {
#pragma unroll
      for (int k = 0; k < D; ++k) out.point[(size_t)i * D + k] = hit ? r.o[k] + r.d[k] * w.res.t : 0.0f;
    }
    if (out.normal) {
      float nv[D];
#pragma unroll
      for (int k = 0; k < D; ++k) nv[k] = 0.0f;
      if (hit) tlas_normal<D, IK>(tv, r, w.res, lim0, nv);
#pragma unroll
      for (int k = 0; k < D; ++k) out.normal[(size_t)i * D + k] = nv[k];
    }
    if (out.sub) out.sub[i] = (hit && w.res.sub != NO_LEAF) ? (int32_t)w.res.sub : -1;
  };
  while (true) {
    bool need = rid == NO_LEAF && !exhausted;
    uint64_t m = __ballot(need);
    while (m) {
      if (next >= end) {
        if (!src.get(next, end)) { exhausted = true; break; }
        continue;
      }
      const uint64_t avail = end - next;
      const uint64_t rank = (uint64_t)__popcll(m & lt_mask);
      if (need && rank < avail) {
        need = false;
        rid = (uint32_t)(next + rank);
        ray_setup<D>(r, q + (size_t)rid * 2 * D);
        lim0 = tmax ? tmax[rid] : max_toi;
        if (!tlas_begin<D>(tv, r, lim0, w)) {
          write(rid);
          rid = NO_LEAF;
        }
      }
      const uint64_t taken = (uint64_t)__popcll(m);
      next += taken < avail ? taken : avail;
      m = __ballot(need);
    }
    if (__ballot(!(exhausted && rid == NO_LEAF)) == 0) break;
    const uint32_t active = (uint32_t)__popcll(__ballot(rid != NO_LEAF));
    if (rid == NO_LEAF) continue;
    // once the pool is exhausted the two-level walk runs without pausing
    // (measured faster there, unlike the single-level walk, where pausing
    // also lets lanes at a leaf test it while others descend)
    if (!tlas_step<D, IK, ANY, EXACT>(tv, r, st, sb, w, WavePause{exhausted ? 0u : pause_below, active})) continue;
    write(rid);
    rid = NO_LEAF;
  }
}

// ---------------------------------------------------------------------------
// Pooled collect (all hits, crossings, overlap, cast-all): one walk per query
// appends the query's entries to a staging area; the host side then scans the
// per-query counts, scatters the entries to their CSR slots and sorts each
// query's list, so the order in which entries were appended does not matter.
// ---------------------------------------------------------------------------

// One appended entry: query (or, from a sub-walk, the sub-walk's position in
// the item blocks), position within that query's (sub-walk's) list, sort key.
struct __align__(16) StageEntry {
  uint32_t q, idx;
  uint64_t key;
};

// Waves claim STAGE_BLK-entry blocks from *blocks and fill them; used[b] is
// the number of entries block b holds.  A block is retired when the next
// append does not fit, so every retired block holds more than STAGE_BLK - 64
// entries.  Blocks at or beyond max_blocks are not written and set *overflow.
constexpr uint32_t STAGE_BLK = 512;
constexpr uint32_t NO_BLOCK = 0xFFFFFFFFu;

struct StageOut {
  StageEntry* e;
  uint32_t* used;
  uint32_t* blocks;
  uint32_t* overflow;
  uint32_t max_blocks;
  uint32_t* counts;        // entries per query
};

// Appends one entry for each calling lane (any subset of the wave).  The
// wave's open block and its fill live in two LDS words that only the lowest
// calling lane reads and writes (volatile: another lane may be that lane
// next time).
struct StageAppend {
  StageOut o;
  volatile uint32_t* slot;     // [0] open block or NO_BLOCK, [1] entries used
  __device__ __forceinline__ void operator()(uint32_t qi, uint32_t idx, uint64_t key) const {
    const uint64_t m = __ballot(1);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t leader = (uint32_t)__builtin_ctzll(m);
    const uint32_t k = (uint32_t)__popcll(m);
    const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    uint32_t blk = 0, used = 0;
    if (lane == leader) {
      blk = slot[0];
      used = slot[1];
      if (blk == NO_BLOCK || used + k > STAGE_BLK) {
        if (blk != NO_BLOCK && blk < o.max_blocks) o.used[blk] = used;
        blk = atomicAdd(o.blocks, 1u);
        used = 0;
        slot[0] = blk;
      }
      slot[1] = used + k;
    }
    blk = (uint32_t)__builtin_amdgcn_readlane((int)blk, (int)leader);
    used = (uint32_t)__builtin_amdgcn_readlane((int)used, (int)leader);
    if (blk < o.max_blocks) o.e[(size_t)blk * STAGE_BLK + used + rank] = StageEntry{qi, idx, key};
    else if (lane == leader) *o.overflow = 1u;
  }
  // Retires the wave's open block (one lane, at the end).
  __device__ __forceinline__ void close() const {
    const uint32_t blk = slot[0];
    if (blk != NO_BLOCK && blk < o.max_blocks) o.used[blk] = slot[1];
  }
};

// Per-query state of a collect walk.  MODE 0: all ray hits, 1: box
// crossings (entry, and exit within max_toi), 2: overlap (touching counts),
// 3: all shape-cast hits.  Keys are hit_key(t, leaf), or the leaf for overlap.
template <int D, int KIND, int MODE, bool FILT>
struct CollectQuery {
  using Pol = typename std::conditional<MODE == 3, CastPolicy<D, KIND>, RayPolicy<D, KIND>>::type;
  static constexpr int QS = MODE == 2 ? 2 * D : (MODE == 3 ? 4 * D : 2 * D);
  static constexpr int PS = Prim<D, KIND>::STRIDE;
  float mn[MODE == 2 ? D : 1], mx[MODE == 2 ? D : 1];
  Pol pol;
  float lim;
  __device__ __forceinline__ bool begin(const SceneView& s, const float* q, uint32_t i, float max_toi, const float* tmax) {
    const float* p = q + (size_t)i * QS;
    if constexpr (MODE == 2) {
#pragma unroll
      for (int k = 0; k < D; ++k) { mn[k] = p[k]; mx[k] = p[D + k]; }
      return box_overlap<D>(s.root, s.root + D, mn, mx);
    } else {
      pol.setup(p);
      lim = tmax ? tmax[i] : max_toi;
      return pol.root(s.root, lim);
    }
  }
  __device__ __forceinline__ void pair(const float* nd, bool& ha, bool& hb) const {
    // the parts of one reference leaf: every primitive of the leaf is tested
    // (a ball or triangle may answer outside its own box)
    if (KIND != KIND_AABB && synthetic_pair<D>(nd)) { ha = hb = true; return; }
    if constexpr (MODE == 2) {
      float amn[D], amx[D], bmn[D], bmx[D];
#pragma unroll
      for (int k = 0; k < D; ++k) { amn[k] = nd[4 * k]; amx[k] = nd[4 * k + 1]; bmn[k] = nd[4 * k + 2]; bmx[k] = nd[4 * k + 3]; }
      ha = box_overlap<D>(amn, amx, mn, mx);
      hb = box_overlap<D>(bmn, bmx, mn, mx);
    } else {
      float ta, tb;
      pol.pair(nd, lim, ha, hb, ta, tb);
    }
  }
  template <class Emit>
  __device__ __forceinline__ void leaf(const SceneView& s, const Filter& f, uint32_t first, uint32_t c, Emit&& emit) const {
    for (uint32_t j = 0; j < c; ++j) {
      const uint32_t pos = first + j;
      const float* p = s.prims + (size_t)pos * PS;
      if constexpr (MODE == 2) {
        if (!overlap_narrow<D, KIND>(p, mn, mx)) continue;
        const uint32_t leaf = s.ids[pos];
        if (FILT && !accepts(f, leaf)) continue;
        emit((uint64_t)leaf);
      } else if constexpr (MODE == 1 && KIND == KIND_AABB) {
        float te, tx;
        int ax;
        if (!slab_exact<D>(pol.r, p, p + D, lim, te, tx, ax)) continue;
        const uint32_t leaf = s.ids[pos];
        if (FILT && !accepts(f, leaf)) continue;
        emit(hit_key(te, leaf));
        if (tx > te) {
          float te2, tx2;
          slab_exact<D>(pol.r, p, p + D, SQ_INF, te2, tx2, ax);
          if (tx2 <= lim) emit(hit_key(tx2, leaf));
        }
      } else {
        float t;
        if (!pol.narrow(p, lim, t)) continue;
        const uint32_t leaf = s.ids[pos];
        if (FILT && !accepts(f, leaf)) continue;
        emit(hit_key(t, leaf));
      }
    }
  }
};

// A query's sub-walk: the walk of query q below node ref (k_collect_expand).
// Sub-walks are written in blocks of ITEM_BLK, claimed one at a time by a
// wave from a block counter; used[b] says how many block b holds.
struct WorkItem {
  uint32_t q, ref;
};
constexpr uint32_t ITEM_BLK = 256;

// Query ranges that are the filled parts of item blocks: a wave takes its own
// block first, then blocks from *counter, skipping empty ones.
struct BlockChunks {
  uint32_t* counter;
  uint64_t n;                  // blocks
  const uint32_t* used;
  uint64_t home, waves;
  bool started = false;
  __device__ __forceinline__ bool get(uint64_t& next, uint64_t& e) {
    while (true) {
      uint64_t c;
      if (!started) { started = true; c = home; }
      else c = waves + (uint64_t)wave_fetch_add(counter);
      if (c >= n) return false;
      const uint32_t u = used[c];
      if (!u) continue;
      next = c * ITEM_BLK;
      e = next + u;
      return true;
    }
  }
};

// One wave walks the ranges src hands it, lanes refilled as their walks
// finish (as pool_walk), each step visiting one leaf range (visit_step) or
// pausing below pause_below descending lanes.  Without ITEMS, position i is
// query i walked from the root, its entries carry (i, index in the query's
// list) and a finished query writes its entry count to o.counts[i].  With
// ITEMS, position i is the sub-walk items[i] (src hands out filled parts of
// item blocks), its entries carry (i, index in the sub-walk's list) and a
// finished sub-walk writes its entry count to item_cnt[i].
// CQ is the per-query state (begin / pair / leaf as CollectQuery).
template <int D, class CQ, bool ITEMS, class Src>
__device__ __forceinline__ void pool_collect_with(const SceneView& s, const float* __restrict__ q, Src src, float max_toi,
                                                  const float* __restrict__ tmax, const Filter& f, const RingStack& st,
                                                  const StageAppend& app, uint32_t pause_below,
                                                  const WorkItem* __restrict__ items, uint32_t* item_cnt) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  const uint32_t root_ref = fbits(s.root[2 * D]);
  CQ cq;
  VisitState v{root_ref, 0};
  uint32_t rid = NO_LEAF, cnt = 0, pos = 0;
  uint64_t next = 0, end = 0;
  bool exhausted = !src.get(next, end);
  while (true) {
    bool need = rid == NO_LEAF && !exhausted;
    uint64_t m = __ballot(need);
    while (m) {
      if (next >= end) {
        if (!src.get(next, end)) { exhausted = true; break; }
        continue;
      }
      const uint64_t avail = end - next;
      const uint64_t rank = (uint64_t)__popcll(m & lt_mask);
      if (need && rank < avail) {
        need = false;
        cnt = 0;
        if constexpr (ITEMS) {
          pos = (uint32_t)(next + rank);
          const WorkItem it = items[pos];
          rid = it.q;
          v = VisitState{it.ref, 0};
          (void)cq.begin(s, q, rid, max_toi, tmax);     // the expand kernel ran the root test
        } else {
          rid = (uint32_t)(next + rank);
          v = VisitState{root_ref, 0};
          if (!cq.begin(s, q, rid, max_toi, tmax)) {
            app.o.counts[rid] = 0;
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
    const bool done = visit_step<D>(
        s, [&](const float* nd, bool& ha, bool& hb) { cq.pair(nd, ha, hb); },
        [&](uint32_t first, uint32_t c) {
          cq.leaf(s, f, first, c, [&](uint64_t key) {
            app(ITEMS ? pos : rid, cnt, key);
            ++cnt;
          });
        },
        st, v, WavePause{pause_below, active});
    if (!done) continue;
    if constexpr (ITEMS) {
      item_cnt[pos] = cnt;
    } else {
      app.o.counts[rid] = cnt;
    }
    rid = NO_LEAF;
  }
  if (lane == 0) app.close();
}

// The built-in kinds and modes.
template <int D, int KIND, int MODE, bool FILT, bool ITEMS, class Src>
__device__ __forceinline__ void pool_collect(const SceneView& s, const float* __restrict__ q, Src src, float max_toi,
                                             const float* __restrict__ tmax, const Filter& f, const RingStack& st,
                                             const StageAppend& app, uint32_t pause_below,
                                             const WorkItem* __restrict__ items, uint32_t* item_cnt) {
  pool_collect_with<D, CollectQuery<D, KIND, MODE, FILT>, ITEMS>(s, q, src, max_toi, tmax, f, st, app, pause_below, items,
                                                                 item_cnt);
}

}  // namespace sqd
