// walk64.h -- the device side of the f64 feature: the pair-record layout, the
// pooled walk and its query policies, shared by f64.hip and by consumer narrow
// phases compiled at run time (rtc.hip), which instantiate the same walk over
// their own leaf test.
//
// Device layout: the reference node array re-packed as pair nodes, both
// children's boxes and refs in one 16-byte-aligned record of 4D + 2 doubles,
//   double box[4D] = a.min[D], a.max[D], b.min[D], b.max[D]
//   uint4  ref     = {ref_a, ref_b, count_a, count_b}   (the last 16 bytes)
// where count 0 marks an interior child (ref = its pair index) and count > 0 a
// leaf range [ref, ref + count) of the leaf-ordered primitives.
#pragma once

#include "pool.h"

namespace sqd {

constexpr int F64_MAX_STACK = 64;   // tree levels the per-lane stack holds (median split: about log2(n) + 1)
constexpr int F64_BLOCK = 128;      // threads per block of the pooled kernels (two waves)
constexpr uint32_t F64_RING = 8;    // stack entries per lane kept in LDS by the pooled kernels (RingStack)
// Leaves of an f64 two-level tree, one record each of R[D][D] (row-major),
// t[D], half[D], world box min[D], max[D]: box instances (two_level.rs:70-113,
// local box [-half, half]) or ball instances (shape_overlap.rs:288-345, a ball
// of radius half[0] at the local origin, local box [-r, r]^D; D = 3).
constexpr int KIND_INST64 = 3;
constexpr int KIND_BALL64 = 4;
template <int KIND> constexpr bool is_inst64() { return KIND == KIND_INST64 || KIND == KIND_BALL64; }
// Leaves of a consumer's f64 narrow phase (run-time compiled): record = world
// box min[D], max[D], then the consumer's doubles.
constexpr int KIND_CONSUMER64 = 5;
// Kinds whose nearest hits report a surface normal (hit_normal).
template <int KIND> constexpr bool has_normals64() { return KIND == KIND_AABB || KIND == KIND_SPHERE; }

struct View64 {
  const double* nodes;      // [np][4D + 2]: boxes, then the refs
  const double* prims;      // leaf-ordered, stride 2D (AABB) or D + 1 (SPHERE)
  const uint32_t* ids;      // original leaf id of each leaf-ordered primitive
  const uint64_t* layers;   // by original leaf id, or null
  double root[12];          // root box: min[D], max[D]
  uint32_t root_ref, root_count;   // root_count > 0: the root is a leaf
  uint32_t n_prims;
  uint32_t stack;           // stack entries a lane can need: the tree's depth
};

__device__ __forceinline__ int64_t total_key64(double f) {
  const int64_t i = __double_as_longlong(f);
  return i ^ (int64_t)(((uint64_t)(i >> 63)) >> 1);
}
// (t, leaf) strictly before (bt, bl) under hit_ordering (query/hit.rs:77-84).
__device__ __forceinline__ bool hit_before64(double t, uint32_t leaf, double bt, uint32_t bl) {
  const int64_t a = total_key64(t), b = total_key64(bt);
  return a < b || (a == b && leaf < bl);
}

template <int D, int KIND> constexpr int stride64() {
  return KIND == KIND_AABB ? 2 * D : KIND == KIND_SPHERE ? D + 1 : D * D + 4 * D;
}

// InstancedGeometry::test_overlap of an instance record: exact for balls (the
// distance from the centre to the region, shape_overlap.rs:59-72,339-341),
// the trait default (true) for boxes (instance.rs:89-98).
template <int D, int KIND>
__device__ __forceinline__ bool inst_overlap64(const double* p, const double* rmn, const double* rmx) {
  if constexpr (KIND == KIND_BALL64) {
    double c[D + 1];
#pragma unroll
    for (int k = 0; k < D; ++k) c[k] = p[D * D + k];
    c[D] = p[D * D + D];
    return overlap_narrow<D, KIND_SPHERE>(c, rmn, rmx);
  } else {
    return true;
  }
}

// Visits every leaf range whose boxes (and every ancestor's) pass node_pass,
// calling leaf(first, count) for each: the walk of the variable-length
// queries, whose limit never shrinks, so the order does not matter.
template <int D, class NodePass, class Leaf>
__device__ __forceinline__ void walk64(const View64& v, NodePass&& node_pass, Leaf&& leaf) {
  if (v.n_prims == 0) return;
  if (!node_pass(v.root, v.root + D)) return;
  if (v.root_count) { leaf(v.root_ref, v.root_count); return; }
  uint32_t stk[F64_MAX_STACK];
  int sp = 0;
  uint32_t node = v.root_ref;
  while (true) {
    const double* b = v.nodes + (size_t)node * (4 * D + 2);
    const uint4 rf = *reinterpret_cast<const uint4*>(b + 4 * D);
    bool ha = node_pass(b, b + D);
    bool hb = node_pass(b + 2 * D, b + 3 * D);
    if (ha && rf.z) { leaf(rf.x, rf.z); ha = false; }
    if (hb && rf.w) { leaf(rf.y, rf.w); hb = false; }
    if (ha && hb) {
      stk[sp++] = rf.y;
      node = rf.x;
    } else if (ha) {
      node = rf.x;
    } else if (hb) {
      node = rf.y;
    } else if (sp > 0) {
      node = stk[--sp];
    } else {
      return;
    }
  }
}

__device__ __forceinline__ bool accepts64(const View64& v, const Filter& f, uint32_t id) {
  if (f.exclude >= 0 && (int64_t)id == f.exclude) return false;
  const uint64_t lay = v.layers ? v.layers[id] : ~0ull;
  return (lay & f.mask) != 0;
}

// Pruning slab of both children of a pair node, branch-free per axis.  It
// never rejects a box the exact slab (slab_exact, aabb.rs:118-148) accepts:
// min/max replace the compare-and-swap, which changes no pass/fail for
// non-parallel axes, and a parallel axis whose origin lies outside rejects
// outright.  Only the pass/fail and the order of the two entries are used;
// primitives are always tested exactly.
template <int D>
__device__ __forceinline__ void slab_pair64(const RayD<D, double>& r, const double* nd, double limit, bool& ha, bool& hb,
                                            double& ta, double& tb) {
  double amin = 0.0, amax = limit, bmin = 0.0, bmax = limit;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const double o = r.o[k];
    if (r.par & (1u << k)) {
      if (o < nd[k] || o > nd[D + k]) { amin = __builtin_huge_val(); amax = -__builtin_huge_val(); }
      if (o < nd[2 * D + k] || o > nd[3 * D + k]) { bmin = __builtin_huge_val(); bmax = -__builtin_huge_val(); }
    } else {
      const double iv = r.inv[k];
      const double a1 = (nd[k] - o) * iv, a2 = (nd[D + k] - o) * iv;
      const double b1 = (nd[2 * D + k] - o) * iv, b2 = (nd[3 * D + k] - o) * iv;
      amin = fmax(amin, fmin(a1, a2)); amax = fmin(amax, fmax(a1, a2));
      bmin = fmax(bmin, fmin(b1, b2)); bmax = fmin(bmax, fmax(b1, b2));
    }
  }
  ha = amin <= amax; hb = bmin <= bmax; ta = amin; tb = bmin;
}

// What a query is for the pooled walk: a ray (RayPolicy's tests in double) or
// a shape cast (CastPolicy<.., double>: the slab against each box grown by the
// probe box, shapecast.rs:70-73, and the kinds' cast leaf tests).
// GUARD: balls and box instances keep a hit only inside the slab interval of
// their world box (SQ_HITS_GUARDED).
template <int D, int KIND, bool CAST, bool GUARD> struct Pol64;
template <int D, int KIND, bool GUARD> struct Pol64<D, KIND, false, GUARD> {
  static constexpr bool CROSS = false;   // one hit per primitive (test_ray)
  static constexpr int QS = 2 * D;   // doubles per query
  RayD<D, double> r;
  __device__ void setup(const double* src) { ray_setup<D>(r, src); }
  __device__ const RayD<D, double>& ray() const { return r; }
  __device__ bool root(const double* rb, double limit) const {
    double te, tx; int ax;
    return slab_exact<D>(r, rb, rb + D, limit, te, tx, ax);
  }
  __device__ void pair(const double* b, double limit, bool& ha, bool& hb, double& ta, double& tb) const {
    slab_pair64<D>(r, b, limit, ha, hb, ta, tb);
  }
  __device__ bool narrow(const double* p, double limit, double& t) const {
    if constexpr (is_inst64<KIND>()) {
      // test_instance (tlas.rs:459-481): the ray in the instance's frame against
      // its local box, or ray_ball about the local origin (with GUARD kept
      // inside the slab interval of its world box)
      double ol[D], dl[D];
      iso_inverse<D>(p, p + D * D, r, ol, dl);
      int ax;
      if constexpr (KIND == KIND_INST64) {
        if (!local_box<D>(ol, dl, p + D * D + D, limit, t, ax)) return false;
      } else {
        if (!ray_ball<D, double>(ol, dl, (const double*)nullptr, p[D * D + D], limit, t)) return false;
      }
      if constexpr (!GUARD) return true;
      double te, tx;
      if (!slab_exact<D>(r, p + D * D + 2 * D, p + D * D + 3 * D, limit, te, tx, ax)) return false;
      return te <= t && t <= tx;
    } else {
      return sqd::narrow<D, KIND, double, GUARD>(r, p, limit, t);
    }
  }
};
template <int D, int KIND, bool GUARD> struct Pol64<D, KIND, true, GUARD> {
  static constexpr bool CROSS = false;
  static constexpr int QS = 4 * D;   // origin, direction, probe box min, max
  // (instance kinds: the cast's slab and probe; the instance test below)
  CastPolicy<D, is_inst64<KIND>() ? KIND_AABB : KIND, double, GUARD> c;
  __device__ void setup(const double* src) { c.setup(src); }
  __device__ const RayD<D, double>& ray() const { return c.r; }
  __device__ bool root(const double* rb, double limit) const { return c.root(rb, limit); }
  __device__ void pair(const double* b, double limit, bool& ha, bool& hb, double& ta, double& tb) const {
    double g[4 * D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      g[k] = b[k] - c.pmax[k]; g[D + k] = b[D + k] - c.pmin[k];
      g[2 * D + k] = b[2 * D + k] - c.pmax[k]; g[3 * D + k] = b[3 * D + k] - c.pmin[k];
    }
    slab_pair64<D>(c.r, g, limit, ha, hb, ta, tb);
  }
  __device__ bool narrow(const double* p, double limit, double& t) const {
    if constexpr (KIND == KIND_INST64) {
      return false;   // InstancedGeometry::test_shape_cast's default: no contact (instance.rs:72-87)
    } else if constexpr (KIND == KIND_BALL64) {
      // shape_overlap.rs:326-337: the cast ray against the ball of radius
      // r + the probe's max x (the built-in SPHERE cast's probe radius) about
      // the centre; GUARD: inside the probe-grown world box's slab interval
      if (!ray_ball<D, double>(c.r.o, c.r.d, p + D * D, p[D * D + D] + c.pmax[0], limit, t)) return false;
      if constexpr (!GUARD) return true;
      double te, tx;
      if (!c.grown_slab(p + D * D + 2 * D, p + D * D + 3 * D, limit, te, tx)) return false;
      return te <= t && t <= tx;
    } else {
      return c.narrow(p, limit, t);
    }
  }
};

// Insertion sort of one query's CSR range [b0, b1): ray and cast hits by
// (toi total order, leaf) as hit_ordering, overlap ids ascending (result lists
// are short).
template <bool OVERLAP>
__device__ void sort_range64(int32_t* out_leaf, double* out_toi, uint64_t b0, uint64_t b1) {
  for (uint64_t a = b0 + 1; a < b1; ++a) {
    const int32_t l = out_leaf[a];
    const double t = OVERLAP ? 0.0 : out_toi[a];
    uint64_t j = a;
    while (j > b0) {
      const int32_t pl = out_leaf[j - 1];
      const bool after = OVERLAP ? (pl > l) : hit_before64(t, (uint32_t)l, out_toi[j - 1], (uint32_t)pl);
      if (!after) break;
      out_leaf[j] = pl;
      if (!OVERLAP) out_toi[j] = out_toi[j - 1];
      --j;
    }
    out_leaf[j] = l;
    if (!OVERLAP) out_toi[j] = t;
  }
}

// Bvh::raycast_nearest (bvh.rs:51-124), the any-hit idiom and
// Bvh::shapecast_nearest (bvh.rs:331-375), pooled; with ALL, Bvh::raycast_all
// (bvh.rs:196-259) and shapecast_all (bvh.rs:378-414): the limit stays the
// query's max_toi and every accepted hit is counted (offsets null: counts[q]
// = hits) or written into the query's CSR range and sorted (offsets given).  Each
// loop iteration refills the wave's idle lanes from the counter (one atomic per
// wave) and advances every walking lane by one pair node.  The wave leaves the
// loop once the counter has passed n and none of its lanes walks; every wave
// then retires once (pool_retire), the last one clearing the counter.  EXACT:
// the reference's hits (unguarded leaf tests; nearest queries over balls and
// box instances in the reference's order, see the REFO step below); else the
// guarded leaf tests near-first.
// Q: the query policy (Pol64, or a consumer's); PS: doubles per primitive record.
template <int D, int KIND, bool ANY, bool CAST, bool ALL, bool EXACT, class Q, int PS>
__device__ __forceinline__ void pool64_walk(const View64& v, const double* __restrict__ rays, uint32_t n, double max_toi,
                                            const double* __restrict__ tmax, const Filter& f, uint32_t* counter,
                                            int32_t* out_leaf, double* out_toi, double* out_point, double* out_normal,
                                            uint8_t* out_any, SpillEntry* spill, uint32_t lanes, uint32_t* counts,
                                            const uint64_t* offsets) {
  const uint32_t lane = threadIdx.x & 63u;
  // the newest F64_RING entries in LDS ([slot][thread] refs, then entries), older ones in this lane's spill column
  extern __shared__ uint32_t f64_ring[];
  const RingStack st{f64_ring + threadIdx.x, reinterpret_cast<float*>(f64_ring + F64_RING * blockDim.x) + threadIdx.x,
                     blockDim.x, F64_RING, spill + (size_t)blockIdx.x * blockDim.x + threadIdx.x, lanes};
  // the reference's order: box leaves (and box casts) hit inside their boxes,
  // so only the other kinds' nearest queries need it
  constexpr bool REFO = EXACT && !ANY && !ALL && KIND != KIND_AABB;
  Q q;
  double best = 0.0, lim = 0.0;   // lim: this query's max_toi
  uint32_t bl = NO_LEAF, bk = 0, node = 0, sp = 0, ray = 0, cnt = 0;
  uint64_t at = 0;   // ALL: the next CSR slot of this query
  bool active = false, drained = false;
  bool resume = false;   // REFO: `node` is a popped parent whose child b is done; its child a comes next
  // leaf range test: the nearest accepted hit so far (ALL: every accepted hit);
  // true ends an any-hit walk
  auto test_leaf = [&](uint32_t first, uint32_t count) {
    for (uint32_t k = first; k < first + count; ++k) {
      const uint32_t id = v.ids[k];
      if (!accepts64(v, f, id)) continue;
      if constexpr (ALL && Q::CROSS) {
        // every crossing the policy reports (test_ray_crossings)
        q.crossings(v.prims + (size_t)k * PS, best, [&](double t) {
          if (offsets) { out_leaf[at] = (int32_t)id; out_toi[at] = t; ++at; }
          ++cnt;
        });
        continue;
      }
      double t;
      if (!q.narrow(v.prims + (size_t)k * PS, best, t)) continue;
      if constexpr (ALL) {
        if (offsets) { out_leaf[at] = (int32_t)id; out_toi[at] = t; ++at; }
        ++cnt;
        continue;
      }
      if (!hit_before64(t, id, best, bl)) continue;
      best = t; bl = id; bk = k;
      if (ANY) return true;
    }
    return false;
  };
  auto finish = [&]() {
    const RayD<D, double>& r = q.ray();
    active = false;
    if constexpr (ALL) {
      if (offsets) sort_range64<false>(out_leaf, out_toi, offsets[ray], offsets[ray + 1]);
      else counts[ray] = cnt;
      return;
    }
    const bool hit = bl != NO_LEAF;
    if (ANY) { out_any[ray] = hit; return; }
    out_leaf[ray] = hit ? (int32_t)bl : -1;
    out_toi[ray] = hit ? best : __builtin_huge_val();
    double nrm[D];
    if constexpr (!CAST && has_normals64<KIND>()) {
      if (hit && out_normal) hit_normal<D, KIND>(r, v.prims + (size_t)bk * PS, best, lim, nrm);
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      // Ray::at (maths/ray.rs), ShapeCast::at for casts (shapecast.rs)
      if (out_point) out_point[(size_t)ray * D + k] = hit ? r.o[k] + r.d[k] * best : 0.0;
      if constexpr (!CAST && has_normals64<KIND>()) {
        if (out_normal) out_normal[(size_t)ray * D + k] = hit ? nrm[k] : 0.0;
      }
    }
  };
  while (true) {
    if (!drained) {
      const uint64_t idle = __ballot(!active);
      if (idle) {
        const int leader = __ffsll((unsigned long long)idle) - 1;
        const uint32_t n_idle = (uint32_t)__popcll(idle);
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(counter, n_idle);
        base = __shfl(base, leader);
        if ((uint64_t)base + n_idle >= n) drained = true;
        if (!active) {
          const uint32_t my = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
          if (my < n) {
            ray = my;
            q.setup(rays + (size_t)my * Q::QS);
            lim = tmax ? tmax[my] : max_toi;
            best = lim; bl = NO_LEAF; sp = 0; resume = false;
            if constexpr (ALL) { cnt = 0; at = offsets ? offsets[my] : 0; }
            const bool in_root = v.n_prims && q.root(v.root, lim);
            if (in_root && !v.root_count) {
              node = v.root_ref;
              active = true;
            } else {
              if (in_root) test_leaf(v.root_ref, v.root_count);
              finish();
            }
          }
        }
      }
    }
    if (drained && !__ballot(active)) break;
    if (!active) continue;
    const double* b = v.nodes + (size_t)node * (4 * D + 2);
    const uint4 rf = *reinterpret_cast<const uint4*>(b + 4 * D);
    double ta, tb;
    bool ha, hb;
    q.pair(b, best, ha, hb, ta, tb);
    if constexpr (REFO) {
      // bvh.rs:73-124: child b's subtree, then child a, each tested at the
      // best when it comes off the reference's stack.  A stack entry is the
      // parent of a pending child a (its stored entry, rounded toward -inf,
      // only screens: the pair is fetched again and a's slab redone at the
      // current best, the reference's test of a when it pops).
      if (resume) { hb = false; resume = false; }
      if (hb && rf.w) {
        test_leaf(rf.y, rf.w);
        hb = false;
        if (ha) {
          double ta2, tb2;
          bool hb2;
          q.pair(b, best, ha, hb2, ta2, tb2);
          ta = ta2;
        }
      }
      if (hb) {
        if (ha) st.push(sp, node, __double2float_rd(ta));
        node = rf.y;
        continue;
      }
      if (ha && !rf.z) { node = rf.x; continue; }
      if (ha) test_leaf(rf.x, rf.z);
      bool found = false;
      while (sp > 0) {
        uint32_t e;
        float lb;
        st.pop(sp, e, lb);
        if (!((double)lb > best)) { node = e; resume = true; found = true; break; }
      }
      if (!found) finish();
      continue;
    }
    bool stop = false;
    if (ha && rf.z) { stop = test_leaf(rf.x, rf.z); ha = false; }
    if (!stop && hb && rf.w) { stop = test_leaf(rf.y, rf.w); hb = false; }
    if (stop) { finish(); continue; }
    if (ha && hb) {
      const bool a_first = ta <= tb;
      // the entry stored rounded toward -inf, so the pop test never prunes a subtree the limit still reaches
      st.push(sp, a_first ? rf.y : rf.x, __double2float_rd(a_first ? tb : ta));
      node = a_first ? rf.x : rf.y;
    } else if (ha) {
      node = rf.x;
    } else if (hb) {
      node = rf.y;
    } else {
      bool found = false;
      while (sp > 0) {
        uint32_t e;
        float lb;
        st.pop(sp, e, lb);
        if (!((double)lb > best)) { node = e; found = true; break; }
      }
      if (!found) finish();
    }
  }
  pool_retire(counter);
}

}  // namespace sqd
