// f64.hip -- the reference's `f64` feature on the device: a Bvh over
// double-precision AABB and SPHERE leaves (D = 2..6) answering nearest-hit,
// any-hit, all-hits, shape-cast (nearest and all) and overlap queries.
//
// Under the feature the reference's Scalar is f64 everywhere
// (crates/spatial-query/src/maths/point.rs:14-18): tree bounds, rays, the slab
// (maths/aabb.rs:118-148), the leaf tests and the returned toi.  The walk
// below runs the same slab and narrow-phase functions as the FP32 engine
// (traverse.h), instantiated on double, over the reference median-split tree
// built in double (host_tree.cpp).  As in the FP32 engine, nearest queries
// over leaves whose hits may fall outside their own box (balls, box instances)
// walk the reference's order (push a, then b, pop LIFO) with the reference's
// unguarded leaf tests, and so answer what the reference's walk does; under
// SQ_HITS_GUARDED those hits are kept only inside their box's slab interval and
// the walk is near-first.
//
// Device layout: the reference node array re-packed as pair nodes, both
// children's boxes and refs in one 16-byte-aligned record of 4D + 2 doubles,
//   double box[4D] = a.min[D], a.max[D], b.min[D], b.max[D]
//   uint4  ref     = {ref_a, ref_b, count_a, count_b}   (the last 16 bytes)
// where count 0 marks an interior child (ref = its pair index) and count > 0 a
// leaf range [ref, ref + count) of the leaf-ordered primitives.  The per-lane
// stack holds at most one entry per tree level.  Nearest and any-hit rays and
// nearest shape casts run pooled: persistent waves whose lanes take the next
// query from an atomic counter as soon as their own ends, so a wave does not
// wait for its longest walk (all hits and all cast contacts too: counted in
// one pass, written in a second); overlap runs one region per lane.
#include "engine_internal.h"
#include "pool.h"

namespace {

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
template <int D, int KIND, bool ANY, bool CAST, bool ALL, bool EXACT>
__global__ __launch_bounds__(F64_BLOCK) void k64_nearest_pool(View64 v, const double* __restrict__ rays, uint32_t n,
                                                        double max_toi, const double* __restrict__ tmax, Filter f,
                                                        uint32_t* counter, int32_t* out_leaf,
                                                        double* out_toi, double* out_point, double* out_normal,
                                                        uint8_t* out_any, SpillEntry* spill, uint32_t lanes,
                                                        uint32_t* counts = nullptr, const uint64_t* offsets = nullptr) {
  constexpr int PS = stride64<D, KIND>();
  const uint32_t lane = threadIdx.x & 63u;
  // the newest F64_RING entries in LDS ([slot][thread] refs, then entries), older ones in this lane's spill column
  extern __shared__ uint32_t f64_ring[];
  const RingStack st{f64_ring + threadIdx.x, reinterpret_cast<float*>(f64_ring + F64_RING * blockDim.x) + threadIdx.x,
                     blockDim.x, F64_RING, spill + (size_t)blockIdx.x * blockDim.x + threadIdx.x, lanes};
  using Q = Pol64<D, KIND, CAST, !EXACT>;
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
    if constexpr (!CAST && !is_inst64<KIND>()) {
      if (hit && out_normal) hit_normal<D, KIND>(r, v.prims + (size_t)bk * PS, best, lim, nrm);
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      // Ray::at (maths/ray.rs), ShapeCast::at for casts (shapecast.rs)
      if (out_point) out_point[(size_t)ray * D + k] = hit ? r.o[k] + r.d[k] * best : 0.0;
      if constexpr (!CAST && !is_inst64<KIND>()) {
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

// Bvh::overlap (bvh.rs:418-456), one query per lane: one pass counts each
// query's leaves, the second writes them into its CSR range and sorts them.
template <int D, int KIND>
__global__ __launch_bounds__(256) void k64_overlap(View64 v, const double* __restrict__ q, uint64_t n, Filter f,
                                                   uint32_t* counts, const uint64_t* offsets, int32_t* out_leaf) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int PS = stride64<D, KIND>();
  const bool fill = offsets != nullptr;
  uint64_t at = fill ? offsets[i] : 0;
  uint32_t cnt = 0;
  const double* rmn = q + i * 2 * D;
  const double* rmx = rmn + D;
  walk64<D>(
      v,
      [&](const double* mn, const double* mx) { return box_overlap<D>(mn, mx, rmn, rmx); },
      [&](uint32_t first, uint32_t count) {
        for (uint32_t k = first; k < first + count; ++k) {
          const uint32_t id = v.ids[k];
          if (!accepts64(v, f, id)) continue;
          if constexpr (is_inst64<KIND>()) {
            if (!inst_overlap64<D, KIND>(v.prims + (size_t)k * PS, rmn, rmx)) continue;
          } else {
            if (!overlap_narrow<D, KIND>(v.prims + (size_t)k * PS, rmn, rmx)) continue;
          }
          if (fill) out_leaf[at++] = (int32_t)id;
          ++cnt;
        }
      });
  if (!fill) { counts[i] = cnt; return; }
  sort_range64<true>(out_leaf, nullptr, offsets[i], offsets[i + 1]);
}

template <class F> int dispatch64(int dim, int kind, F&& f) {
  if (kind != KIND_AABB && kind != KIND_SPHERE)
    return fail(SQ_ERR_UNSUPPORTED, "f64 trees hold AABB or SPHERE primitives");
#define SQ_CASE(DD)                                                       \
  case DD:                                                                \
    return kind == KIND_AABB ? f.template operator()<DD, KIND_AABB>()    \
                             : f.template operator()<DD, KIND_SPHERE>();
  switch (dim) {
    SQ_CASE(2) SQ_CASE(3) SQ_CASE(4) SQ_CASE(5) SQ_CASE(6)
    default: return fail(SQ_ERR_UNSUPPORTED, "dimension " + std::to_string(dim) + " not instantiated (2..6)");
  }
#undef SQ_CASE
}

// The instance kinds of an f64 Tlas: boxes in D = 2..6, balls in D = 3.
template <class F> int dispatch_inst64(int dim, int kind, F&& f) {
  if (kind == KIND_BALL64) {
    if (dim != 3) return fail(SQ_ERR_UNSUPPORTED, "ball instances need dim 3");
    return f.template operator()<3, KIND_BALL64>();
  }
  switch (dim) {
    case 2: return f.template operator()<2, KIND_INST64>();
    case 3: return f.template operator()<3, KIND_INST64>();
    case 4: return f.template operator()<4, KIND_INST64>();
    case 5: return f.template operator()<5, KIND_INST64>();
    case 6: return f.template operator()<6, KIND_INST64>();
    default: return fail(SQ_ERR_UNSUPPORTED, "dimension " + std::to_string(dim) + " not instantiated (2..6)");
  }
}

}  // namespace

struct sq_bvh64 {
  sq_ctx* ctx = nullptr;
  int dim = 0, kind = 0;
  uint64_t n_prims = 0;
  sqh::RefTreeT<double> tree;
  DevBuf nodes, prims, ids, layers;
  View64 view{};
  sq_tree_stats stats{};
};

namespace {

Filter filter64(const sq_bvh64* b, const sq_filter* f) {
  Filter fd;
  fd.layers = b->layers.p ? b->layers.as<uint64_t>() : nullptr;
  fd.mask = f ? f->layer_mask : ~0ull;
  fd.exclude = f ? f->exclude : -1;
  return fd;
}

}  // namespace

// One pooled launch over n queries (the grid fills the chip; lanes refill from
// the context's pool counter, which every launch leaves zero).
template <int D, int KIND, bool ANY, bool CAST, bool ALL, bool EXACT = true>
static int pool_launch(sq_bvh64* b, const double* q, uint64_t n, double max_toi, const double* tmax, const Filter& fd,
                       int32_t* out_leaf, double* out_toi, double* out_point, double* out_normal, uint8_t* out_any,
                       uint32_t* counts, const uint64_t* offsets) {
  sq_ctx* c = b->ctx;
  const void* kern = reinterpret_cast<const void*>(&k64_nearest_pool<D, KIND, ANY, CAST, ALL, EXACT>);
  const size_t lds = (size_t)F64_RING * F64_BLOCK * 8;
  int per_cu = 0;
  if (int r0 = blocks_per_cu(kern, c->device, F64_BLOCK, lds, &per_cu)) return r0;
  const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)c->cus * std::max(1, per_cu), grid_for(n, F64_BLOCK)));
  const uint64_t lanes = grid * F64_BLOCK;
  // spill rows: the stack holds at most depth - 2 entries, F64_RING of them in LDS
  const uint64_t rows = b->view.stack > F64_RING ? b->view.stack - F64_RING : 0;
  if (int r1 = c->spill.ensure((size_t)rows * lanes * sizeof(SpillEntry) + 64)) return r1;
  k64_nearest_pool<D, KIND, ANY, CAST, ALL, EXACT><<<(uint32_t)grid, F64_BLOCK, lds, c->stream>>>(
      b->view, q, (uint32_t)n, max_toi, tmax, fd, c->pool_ctr.as<uint32_t>(), out_leaf, out_toi, out_point, out_normal,
      out_any, c->spill.as<SpillEntry>(), (uint32_t)lanes, counts, offsets);
  HIP_TRY(hipGetLastError());
  return SQ_OK;
}

extern "C" {

// The f64 tree over n leaves with world boxes `boxes` [n][2D] and device
// records `prims` [n][ps] (leaf-ordered on upload).
static int create64(sq_ctx* c, int dim, int kind, const double* prims, int ps, const std::vector<double>& boxes, uint64_t n,
                    sq_bvh64** out) {
  int rc;
  std::unique_ptr<sq_bvh64> b(new sq_bvh64());
  b->ctx = c; b->dim = dim; b->kind = kind; b->n_prims = n;
  sqh::build_tree(dim, boxes.data(), (uint32_t)n, b->tree);
  const sqh::RefTreeT<double>& t = b->tree;
  uint64_t leaves = 0, nprims = 0, depth = 0;
  sqh::tree_structural(t, &leaves, &nprims, &depth);
  if (depth > F64_MAX_STACK) return fail(SQ_ERR_UNSUPPORTED, "tree deeper than the f64 walk's stack");
  b->view.stack = (uint32_t)depth;
  // TreeStats (tree.rs:81-110, bvh.rs:477-489) of the f64 tree: size_of::<Node<D>>() is 16D + 12 bytes
  // padded to the f64 alignment; the interior surface sum is summed in double and reported as float
  b->stats.node_count = t.node_count();
  b->stats.leaf_count = leaves;
  b->stats.prim_count = nprims;
  b->stats.max_depth = depth;
  b->stats.bytes = b->stats.node_count * (uint64_t)(16 * dim + 16) + t.prim_indices.size() * 4;
  double quality = 0.0;
  for (size_t i = 0; i < t.node_count(); ++i) {
    if (t.prim_count[i] > 0) continue;
    const double* bx = t.bounds.data() + i * 2 * dim;
    double sum = 0.0;   // Aabb::half_surface, aabb.rs:84-97
    for (int a = 0; a < dim; ++a) {
      double prod = 1.0;
      for (int k = 0; k < dim; ++k) if (k != a) prod = prod * (bx[dim + k] - bx[k]);
      sum = sum + prod;
    }
    quality = quality + sum;
  }
  b->stats.quality = (float)quality;
  // pair nodes in reference pre-order of their interior nodes
  const size_t nn = t.node_count();
  std::vector<uint32_t> pair_of(nn, 0);
  uint32_t np = 0;
  for (size_t i = 0; i < nn; ++i) if (t.prim_count[i] == 0) pair_of[i] = np++;
  const size_t ns = 4 * (size_t)dim + 2;   // doubles per pair record
  std::vector<double> pb((size_t)np * ns);
  for (size_t i = 0; i < nn; ++i) {
    if (t.prim_count[i] != 0) continue;
    const uint32_t p = pair_of[i];
    const uint32_t ch[2] = {t.child_a[i], t.child_b[i]};
    uint32_t ref[4];
    for (int s = 0; s < 2; ++s) {
      const uint32_t cn = ch[s];
      std::memcpy(pb.data() + (size_t)p * ns + 2 * s * dim, t.bounds.data() + (size_t)cn * 2 * dim, 2 * dim * sizeof(double));
      const bool leaf = t.prim_count[cn] != 0;
      ref[s] = leaf ? t.child_a[cn] : pair_of[cn];
      ref[2 + s] = leaf ? t.prim_count[cn] : 0;
    }
    std::memcpy(pb.data() + (size_t)p * ns + 4 * dim, ref, sizeof(ref));
  }
  // leaf-ordered primitives and their original ids
  std::vector<double> lp((size_t)n * ps);
  std::vector<uint32_t> ids(n);
  for (uint64_t k = 0; k < n; ++k) {
    const uint32_t id = t.prim_indices[k];
    std::memcpy(lp.data() + k * ps, prims + (size_t)id * ps, ps * sizeof(double));
    ids[k] = id;
  }
  HIP_TRY(hipSetDevice(c->device));
  if ((rc = b->nodes.ensure(pb.size() * 8)) || (rc = b->prims.ensure(lp.size() * 8)) ||
      (rc = b->ids.ensure(ids.size() * 4)))
    return rc;
  if (np) HIP_TRY(hipMemcpy(b->nodes.p, pb.data(), pb.size() * 8, hipMemcpyHostToDevice));
  if (n) {
    HIP_TRY(hipMemcpy(b->prims.p, lp.data(), lp.size() * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->ids.p, ids.data(), ids.size() * 4, hipMemcpyHostToDevice));
  }
  View64& v = b->view;
  v.nodes = b->nodes.as<double>();
  v.prims = b->prims.as<double>();
  v.ids = b->ids.as<uint32_t>();
  v.layers = nullptr;
  v.n_prims = (uint32_t)n;
  if (nn) {
    std::memcpy(v.root, t.bounds.data(), 2 * dim * sizeof(double));
    v.root_count = t.prim_count[0];
    v.root_ref = t.prim_count[0] ? t.child_a[0] : pair_of[0];
  }
  ctx_retain(c);
  *out = b.release();
  return SQ_OK;
}

int sq_bvh64_create(sq_ctx* c, int dim, int kind, const double* prims, uint64_t n, sq_bvh64** out) {
  if (!c || !out || (n && !prims)) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (n >= 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "too many primitives (ids are u32)");
  int rc = dispatch64(dim, kind, [&]<int D, int KIND>() -> int { return SQ_OK; });
  if (rc) return rc;
  const int ps = kind == KIND_AABB ? 2 * dim : dim + 1;
  // QueryGeometry::world_aabb of the built-in kinds, in double
  std::vector<double> boxes((size_t)n * 2 * dim);
  for (uint64_t i = 0; i < n; ++i) {
    const double* p = prims + i * ps;
    double* bx = boxes.data() + i * 2 * dim;
    for (int k = 0; k < dim; ++k) {
      if (kind == KIND_AABB) { bx[k] = p[k]; bx[dim + k] = p[dim + k]; }
      else { bx[k] = p[k] - p[dim]; bx[dim + k] = p[k] + p[dim]; }
    }
    if (kind == KIND_SPHERE && !(p[dim] >= 0.0)) return fail(SQ_ERR_INVALID, "sphere radius must be >= 0");
  }
  return create64(c, dim, kind, prims, ps, boxes, n, out);
}

// Tlas::build (tlas.rs:35-56) over box or ball instances in double: each world
// box is Isometry::transform_aabb of the local box, [-half, half] or
// [-r, r]^D (isometry.rs:84-97).  Balls store splat(r) as their half extents.
int sq_tlas64_create_kind(sq_ctx* c, int dim, int inst_kind, uint64_t n, const double* rot, const double* trans,
                          const double* halves, sq_tlas64** out) {
  if (!c || !out || (n && (!rot || !trans || !halves))) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (inst_kind != SQ_INST_BOX && inst_kind != SQ_INST_BALL)
    return fail(SQ_ERR_UNSUPPORTED, "f64 instance trees hold SQ_INST_BOX or SQ_INST_BALL instances");
  const bool ball = inst_kind == SQ_INST_BALL;
  if (ball && dim != 3) return fail(SQ_ERR_UNSUPPORTED, "ball instances need dim 3");
  if (dim < 2 || dim > 6) return fail(SQ_ERR_UNSUPPORTED, "dimension " + std::to_string(dim) + " not instantiated (2..6)");
  if (n >= 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "too many instances (ids are u32)");
  const int D = dim, ps = D * D + 4 * D;
  std::vector<double> boxes((size_t)n * 2 * D), recs((size_t)n * ps);
  for (uint64_t i = 0; i < n; ++i) {
    const double* R = rot + i * D * D;
    const double* t = trans + i * D;
    double h[6];
    for (int k = 0; k < D; ++k) h[k] = ball ? halves[i] : halves[i * D + k];
    double* rec = recs.data() + i * ps;
    for (int k = 0; k < D; ++k)
      if (!(h[k] >= 0.0)) return fail(SQ_ERR_INVALID, ball ? "ball radii must be >= 0" : "instance half extents must be >= 0");
    std::memcpy(rec, R, D * D * sizeof(double));
    std::memcpy(rec + D * D, t, D * sizeof(double));
    std::memcpy(rec + D * D + D, h, D * sizeof(double));
    // transform_aabb: centre R c + t (c = 0 for [-h, h]), half |R| h
    double* w = boxes.data() + i * 2 * D;
    for (int a = 0; a < D; ++a) {
      double cc = 0.0;
      for (int b = 0; b < D; ++b) cc = cc + R[a * D + b] * ((h[b] * -1.0 + h[b]) * 0.5);
      const double rc = cc + t[a];
      double acc = 0.0;
      for (int b = 0; b < D; ++b) acc = acc + std::fabs(R[a * D + b]) * ((h[b] - h[b] * -1.0) * 0.5);
      w[a] = rc - acc;
      w[D + a] = rc + acc;
    }
    std::memcpy(rec + D * D + 2 * D, w, 2 * D * sizeof(double));
  }
  sq_bvh64* b = nullptr;
  if (int rc = create64(c, dim, ball ? KIND_BALL64 : KIND_INST64, recs.data(), ps, boxes, n, &b)) return rc;
  *out = reinterpret_cast<sq_tlas64*>(b);
  return SQ_OK;
}

int sq_tlas64_create(sq_ctx* c, int dim, uint64_t n, const double* rot, const double* trans, const double* halves,
                     sq_tlas64** out) {
  return sq_tlas64_create_kind(c, dim, SQ_INST_BOX, n, rot, trans, halves, out);
}

int sq_tlas64_destroy(sq_tlas64* t) { return sq_bvh64_destroy(reinterpret_cast<sq_bvh64*>(t)); }

int sq_tlas64_stats(const sq_tlas64* t, sq_tree_stats* out) {
  return sq_bvh64_stats(reinterpret_cast<const sq_bvh64*>(t), out);
}

int sq_bvh64_destroy(sq_bvh64* b) {
  if (!b) return SQ_OK;
  sq_ctx* c = b->ctx;
  if (c) (void)hipStreamSynchronize(c->stream);
  delete b;
  if (c) ctx_release(c);
  return SQ_OK;
}

int sq_bvh64_stats(const sq_bvh64* b, sq_tree_stats* out) {
  if (!b || !out) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = b->stats;
  return SQ_OK;
}

int sq_bvh64_set_layers(sq_bvh64* b, const uint64_t* layers) {
  if (!b) return fail(SQ_ERR_INVALID, "bvh is NULL");
  if (!layers) {
    b->layers.~DevBuf();
    new (&b->layers) DevBuf();
  } else {
    int rc = b->layers.ensure(std::max<uint64_t>(b->n_prims, 1) * 8);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(b->layers.p, layers, b->n_prims * 8, hipMemcpyDefault, b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));   // the caller may free a host array on return
  }
  b->view.layers = b->layers.p ? b->layers.as<uint64_t>() : nullptr;
  return SQ_OK;
}

static int nearest64(sq_bvh64* b, const double* rays, uint64_t n, double max_toi, const double* tmax, const sq_filter* f,
                     int32_t* out_leaf, double* out_toi, double* out_point, double* out_normal, uint8_t* out_any,
                     bool cast) {
  if (!b || (n && !rays)) return fail(SQ_ERR_INVALID, "NULL argument");
  if (n && !out_any && (!out_leaf || !out_toi)) return fail(SQ_ERR_INVALID, "output arrays are NULL");
  if (!n) return SQ_OK;
  if (n >= (1ull << 31)) return fail(SQ_ERR_INVALID, "more than 2^31 queries in one call");
  const Filter fd = filter64(b, f);
  const bool guarded = b->ctx->hit_mode == SQ_HITS_GUARDED;
  if (b->kind == KIND_INST64 || b->kind == KIND_BALL64) {
    return dispatch_inst64(b->dim, b->kind, [&]<int D, int IK>() -> int {
      auto run = [&]<bool EXACT>() -> int {
        if (cast)
          return pool_launch<D, IK, false, true, false, EXACT>(b, rays, n, max_toi, tmax, fd, out_leaf, out_toi, out_point,
                                                               nullptr, nullptr, nullptr, nullptr);
        if (out_any)
          return pool_launch<D, IK, true, false, false, EXACT>(b, rays, n, max_toi, tmax, fd, nullptr, nullptr, nullptr,
                                                               nullptr, out_any, nullptr, nullptr);
        return pool_launch<D, IK, false, false, false, EXACT>(b, rays, n, max_toi, tmax, fd, out_leaf, out_toi, out_point,
                                                              nullptr, nullptr, nullptr, nullptr);
      };
      return guarded ? run.template operator()<false>() : run.template operator()<true>();
    });
  }
  return dispatch64(b->dim, b->kind, [&]<int D, int KIND>() -> int {
    auto run = [&]<bool EXACT>() -> int {
      if (cast)
        return pool_launch<D, KIND, false, true, false, EXACT>(b, rays, n, max_toi, tmax, fd, out_leaf, out_toi, out_point,
                                                               nullptr, nullptr, nullptr, nullptr);
      if (out_any)
        return pool_launch<D, KIND, true, false, false, EXACT>(b, rays, n, max_toi, tmax, fd, nullptr, nullptr, nullptr,
                                                               nullptr, out_any, nullptr, nullptr);
      return pool_launch<D, KIND, false, false, false, EXACT>(b, rays, n, max_toi, tmax, fd, out_leaf, out_toi, out_point,
                                                              out_normal, nullptr, nullptr, nullptr);
    };
    // box leaves hit inside their boxes: the one walk is exact in both modes
    if constexpr (KIND == KIND_AABB) return run.template operator()<true>();
    else return guarded ? run.template operator()<false>() : run.template operator()<true>();
  });
}

int sq_bvh64_raycast_nearest(sq_bvh64* b, const double* rays, uint64_t n, double max_toi, const double* max_toi_per_ray,
                             const sq_filter* f, int32_t* out_leaf, double* out_toi, double* out_point,
                             double* out_normal) {
  return nearest64(b, rays, n, max_toi, max_toi_per_ray, f, out_leaf, out_toi, out_point, out_normal, nullptr, false);
}

int sq_bvh64_shapecast_nearest(sq_bvh64* b, const double* casts, uint64_t n, double max_toi,
                               const double* max_toi_per_cast, const sq_filter* f, int32_t* out_leaf, double* out_toi,
                               double* out_point) {
  return nearest64(b, casts, n, max_toi, max_toi_per_cast, f, out_leaf, out_toi, out_point, nullptr, nullptr, true);
}

int sq_bvh64_raycast_any(sq_bvh64* b, const double* rays, uint64_t n, double max_toi, const double* max_toi_per_ray,
                         const sq_filter* f, uint8_t* out_hit) {
  if (n && !out_hit) return fail(SQ_ERR_INVALID, "NULL argument");
  return nearest64(b, rays, n, max_toi, max_toi_per_ray, f, nullptr, nullptr, nullptr, nullptr, out_hit, false);
}

int sq_tlas64_raycast_nearest(sq_tlas64* t, const double* rays, uint64_t n, double max_toi,
                              const double* max_toi_per_ray, int32_t* out_inst, double* out_toi, double* out_point) {
  return nearest64(reinterpret_cast<sq_bvh64*>(t), rays, n, max_toi, max_toi_per_ray, nullptr, out_inst, out_toi,
                   out_point, nullptr, nullptr, false);
}

int sq_tlas64_raycast_any(sq_tlas64* t, const double* rays, uint64_t n, double max_toi, const double* max_toi_per_ray,
                          uint8_t* out_hit) {
  if (n && !out_hit) return fail(SQ_ERR_INVALID, "NULL argument");
  return nearest64(reinterpret_cast<sq_bvh64*>(t), rays, n, max_toi, max_toi_per_ray, nullptr, nullptr, nullptr, nullptr,
                   nullptr, out_hit, false);
}

// mode 0: rays, 1: overlap regions, 2: shape casts
static int collect64(sq_bvh64* b, const double* q, uint64_t n, double max_toi, const double* tmax, const sq_filter* f,
                     int mode,
                     uint64_t* offsets, int32_t* out_leaf, double* out_toi, uint64_t capacity, uint64_t* total) {
  const bool overlap = mode == 1;
  if (!b || !offsets || !total || (n && !q)) return fail(SQ_ERR_INVALID, "NULL argument");
  sq_ctx* c = b->ctx;
  *total = 0;
  if (!n) {
    HIP_TRY(hipMemsetAsync(offsets, 0, 8, c->stream));
    return SQ_OK;
  }
  if (n >= (1ull << 31)) return fail(SQ_ERR_INVALID, "more than 2^31 queries in one call");
  int rc;
  if ((rc = c->counts.ensure(n * 4))) return rc;
  const Filter fd = filter64(b, f);
  uint32_t* counts = c->counts.as<uint32_t>();
  // rays and casts: the pooled walk (count pass, then fill pass); overlap: one region per lane
  // (lists never shrink a limit: every kind keeps its unguarded hits in both modes)
  auto body = [&]<int D, int KIND>(const uint64_t* offs) -> int {
    if (mode == 1) {
      k64_overlap<D, KIND><<<grid_for(n, 256), 256, 0, c->stream>>>(b->view, q, n, fd, counts, offs, out_leaf);
      HIP_TRY(hipGetLastError());
      return SQ_OK;
    }
    if (mode == 2)
      return pool_launch<D, KIND, false, true, true>(b, q, n, max_toi, tmax, fd, out_leaf, out_toi, nullptr, nullptr,
                                                     nullptr, counts, offs);
    return pool_launch<D, KIND, false, false, true>(b, q, n, max_toi, tmax, fd, out_leaf, out_toi, nullptr, nullptr,
                                                    nullptr, counts, offs);
  };
  auto launch = [&](const uint64_t* offs) -> int {
    if (b->kind == KIND_INST64 || b->kind == KIND_BALL64)
      return dispatch_inst64(b->dim, b->kind, [&]<int D, int KIND>() -> int { return body.template operator()<D, KIND>(offs); });
    return dispatch64(b->dim, b->kind, [&]<int D, int KIND>() -> int { return body.template operator()<D, KIND>(offs); });
  };
  if ((rc = launch(nullptr))) return rc;
  std::vector<std::unique_ptr<DevBuf>> temps;
  if ((rc = scan_counts(c, c->counts.as<uint32_t>(), n, offsets, temps))) return rc;
  HIP_TRY(hipMemcpyAsync(total, offsets + n, 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const uint64_t tot = *total;
  if (tot > capacity)
    return fail(SQ_ERR_CAPACITY, "result needs " + std::to_string(tot) + " entries, capacity is " + std::to_string(capacity));
  if (!tot) return SQ_OK;
  if (!out_leaf || (!overlap && !out_toi)) return fail(SQ_ERR_INVALID, "output arrays are NULL");
  return launch(offsets);
}

int sq_bvh64_raycast_all(sq_bvh64* b, const double* rays, uint64_t n, double max_toi, const sq_filter* f,
                         uint64_t* offsets, int32_t* out_leaf, double* out_toi, uint64_t capacity, uint64_t* total) {
  return collect64(b, rays, n, max_toi, nullptr, f, 0, offsets, out_leaf, out_toi, capacity, total);
}

int sq_bvh64_shapecast_all(sq_bvh64* b, const double* casts, uint64_t n, double max_toi, const double* max_toi_per_cast,
                           const sq_filter* f, uint64_t* offsets, int32_t* out_leaf, double* out_toi, uint64_t capacity,
                           uint64_t* total) {
  return collect64(b, casts, n, max_toi, max_toi_per_cast, f, 2, offsets, out_leaf, out_toi, capacity, total);
}

int sq_bvh64_overlap(sq_bvh64* b, const double* regions, uint64_t n, const sq_filter* f, uint64_t* offsets,
                     int32_t* out_leaf, uint64_t capacity, uint64_t* total) {
  return collect64(b, regions, n, 0.0, nullptr, f, 1, offsets, out_leaf, nullptr, capacity, total);
}

// Tlas::raycast_all (tlas.rs:158-215), also Tlas::raycast_crossings
// (tlas.rs:221-304) for these kinds: one crossing per instance, the
// test_ray_crossings_local default.
int sq_tlas64_raycast_all(sq_tlas64* t, const double* rays, uint64_t n, double max_toi, uint64_t* offsets,
                          int32_t* out_inst, double* out_toi, uint64_t capacity, uint64_t* total) {
  return collect64(reinterpret_cast<sq_bvh64*>(t), rays, n, max_toi, nullptr, nullptr, 0, offsets, out_inst, out_toi,
                   capacity, total);
}

// Tlas::overlap (tlas.rs:393-432): every instance of a leaf node touching the
// region that passes test_overlap (exact for balls, true for boxes).
int sq_tlas64_overlap(sq_tlas64* t, const double* regions, uint64_t n, uint64_t* offsets, int32_t* out_inst,
                      uint64_t capacity, uint64_t* total) {
  return collect64(reinterpret_cast<sq_bvh64*>(t), regions, n, 0.0, nullptr, nullptr, 1, offsets, out_inst, nullptr,
                   capacity, total);
}

// Tlas::shapecast_nearest (tlas.rs:306-351) and shapecast_all (:353-391).
int sq_tlas64_shapecast_nearest(sq_tlas64* t, const double* casts, uint64_t n, double max_toi, int32_t* out_inst,
                                double* out_toi, double* out_point) {
  return nearest64(reinterpret_cast<sq_bvh64*>(t), casts, n, max_toi, nullptr, nullptr, out_inst, out_toi, out_point,
                   nullptr, nullptr, true);
}

int sq_tlas64_shapecast_all(sq_tlas64* t, const double* casts, uint64_t n, double max_toi, uint64_t* offsets,
                            int32_t* out_inst, double* out_toi, uint64_t capacity, uint64_t* total) {
  return collect64(reinterpret_cast<sq_bvh64*>(t), casts, n, max_toi, nullptr, nullptr, 2, offsets, out_inst, out_toi,
                   capacity, total);
}

}  // extern "C"
