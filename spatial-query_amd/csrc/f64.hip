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
// The device layout (pair records of 4D + 2 doubles) and the pooled walk are
// in walk64.h, which run-time compiled consumer narrow phases share.  The per-lane
// stack holds at most one entry per tree level.  Nearest and any-hit rays and
// nearest shape casts run pooled: persistent waves whose lanes take the next
// query from an atomic counter as soon as their own ends, so a wave does not
// wait for its longest walk (all hits and all cast contacts too: counted in
// one pass, written in a second); overlap runs one region per lane.
#include "engine_internal.h"
#include "walk64.h"

namespace {

// The pooled walk over the built-in f64 kinds (walk64.h pool64_walk with
// Pol64): nearest, any-hit, all hits, nearest and all shape casts.
template <int D, int KIND, bool ANY, bool CAST, bool ALL, bool EXACT>
__global__ __launch_bounds__(F64_BLOCK) void k64_nearest_pool(View64 v, const double* __restrict__ rays, uint32_t n,
                                                        double max_toi, const double* __restrict__ tmax, Filter f,
                                                        uint32_t* counter, int32_t* out_leaf,
                                                        double* out_toi, double* out_point, double* out_normal,
                                                        uint8_t* out_any, SpillEntry* spill, uint32_t lanes,
                                                        uint32_t* counts = nullptr, const uint64_t* offsets = nullptr) {
  pool64_walk<D, KIND, ANY, CAST, ALL, EXACT, Pol64<D, KIND, CAST, !EXACT>, stride64<D, KIND>()>(
      v, rays, n, max_toi, tmax, f, counter, out_leaf, out_toi, out_point, out_normal, out_any, spill, lanes, counts, offsets);
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
  if (kind == KIND_CONSUMER64)
    return fail(SQ_ERR_INVALID, "a consumer tree answers through its narrow phase (sq_raycast64_*_consumer)");
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
  uint32_t consumer_pd = 0;   // KIND_CONSUMER64: the narrow phase's doubles per primitive
  int ps = 0;                 // doubles per leaf record
  double baseline = 0.0;      // interior surface sum at the build (Tlas::needs_rebuild)
  std::vector<double> inst_half;   // instance trees: the half extents (or radii) last given
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

// tree::interior_surface_sum (tree.rs:104-110) of the f64 tree: the half
// surfaces (Aabb::half_surface, aabb.rs:84-97) of the interior nodes summed in
// node order, in double.
static double surface_sum64(const sqh::RefTreeT<double>& t) {
  const int dim = t.dim;
  double quality = 0.0;
  for (size_t i = 0; i < t.node_count(); ++i) {
    if (t.prim_count[i] > 0) continue;
    const double* bx = t.bounds.data() + i * 2 * dim;
    double sum = 0.0;
    for (int a = 0; a < dim; ++a) {
      double prod = 1.0;
      for (int k = 0; k < dim; ++k) if (k != a) prod = prod * (bx[dim + k] - bx[k]);
      sum = sum + prod;
    }
    quality = quality + sum;
  }
  return quality;
}

// tree::refit (tree.rs:54-75) over fresh leaf boxes [n][2D]: one reverse pass,
// a leaf folding its primitives' boxes from EMPTY in prim_indices order, an
// interior node the union of its children (componentwise min and max, exact).
static void refit_bounds64(sqh::RefTreeT<double>& t, const std::vector<double>& boxes) {
  const int D = t.dim;
  for (size_t ni = t.node_count(); ni-- > 0;) {
    double* b = t.bounds.data() + ni * 2 * D;
    if (t.prim_count[ni] > 0) {
      for (int k = 0; k < D; ++k) { b[k] = INFINITY; b[D + k] = -INFINITY; }
      for (uint32_t j = t.child_a[ni]; j < t.child_a[ni] + t.prim_count[ni]; ++j) {
        const double* pb = boxes.data() + (size_t)t.prim_indices[j] * 2 * D;
        for (int k = 0; k < D; ++k) { b[k] = std::fmin(b[k], pb[k]); b[D + k] = std::fmax(b[D + k], pb[D + k]); }
      }
    } else {
      const double* ba = t.bounds.data() + (size_t)t.child_a[ni] * 2 * D;
      const double* bb = t.bounds.data() + (size_t)t.child_b[ni] * 2 * D;
      for (int k = 0; k < D; ++k) { b[k] = std::fmin(ba[k], bb[k]); b[D + k] = std::fmax(ba[D + k], bb[D + k]); }
    }
  }
}

// The pair records, leaf-ordered records `prims` [n][ps] (caller order) and
// ids of b->tree to the device, and the root bounds into the walk's view.  A
// refit uploads the same sizes again: the arrays stay where they are.
static int upload64(sq_bvh64* b, const double* prims) {
  sq_ctx* c = b->ctx;
  const int dim = b->dim, ps = b->ps;
  const uint64_t n = b->n_prims;
  const sqh::RefTreeT<double>& t = b->tree;
  int rc;
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
  // (walks already enqueued on the context stream read the old arrays)
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (np) HIP_TRY(hipMemcpy(b->nodes.p, pb.data(), pb.size() * 8, hipMemcpyHostToDevice));
  if (n) {
    HIP_TRY(hipMemcpy(b->prims.p, lp.data(), lp.size() * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->ids.p, ids.data(), ids.size() * 4, hipMemcpyHostToDevice));
  }
  View64& v = b->view;
  v.nodes = b->nodes.as<double>();
  v.prims = b->prims.as<double>();
  v.ids = b->ids.as<uint32_t>();
  v.n_prims = (uint32_t)n;
  if (nn) {
    std::memcpy(v.root, t.bounds.data(), 2 * dim * sizeof(double));
    v.root_count = t.prim_count[0];
    v.root_ref = t.prim_count[0] ? t.child_a[0] : pair_of[0];
  }
  b->stats.quality = (float)surface_sum64(t);
  return SQ_OK;
}

// The f64 tree over n leaves with world boxes `boxes` [n][2D] and device
// records `prims` [n][ps] (leaf-ordered on upload).
static int create64(sq_ctx* c, int dim, int kind, const double* prims, int ps, const std::vector<double>& boxes, uint64_t n,
                    sq_bvh64** out) {
  int rc;
  std::unique_ptr<sq_bvh64> b(new sq_bvh64());
  b->ctx = c; b->dim = dim; b->kind = kind; b->n_prims = n; b->ps = ps;
  sqh::build_tree(dim, boxes.data(), (uint32_t)n, b->tree);
  const sqh::RefTreeT<double>& t = b->tree;
  uint64_t leaves = 0, nprims = 0, depth = 0;
  sqh::tree_structural(t, &leaves, &nprims, &depth);
  if (depth > F64_MAX_STACK) return fail(SQ_ERR_UNSUPPORTED, "tree deeper than the f64 walk's stack");
  b->view.stack = (uint32_t)depth;
  b->view.layers = nullptr;
  // TreeStats (tree.rs:81-110, bvh.rs:477-489) of the f64 tree: size_of::<Node<D>>() is 16D + 12 bytes
  // padded to the f64 alignment; the interior surface sum is summed in double and reported as float
  b->stats.node_count = t.node_count();
  b->stats.leaf_count = leaves;
  b->stats.prim_count = nprims;
  b->stats.max_depth = depth;
  b->stats.bytes = b->stats.node_count * (uint64_t)(16 * dim + 16) + t.prim_indices.size() * 4;
  b->baseline = surface_sum64(t);
  if ((rc = upload64(b.get(), prims))) return rc;
  ctx_retain(c);
  *out = b.release();
  return SQ_OK;
}

// QueryGeometry::world_aabb of the built-in kinds, in double: the box, or
// Aabb::new(c - splat(r), c + splat(r)).
static int prim_boxes64(int dim, int kind, const double* prims, uint64_t n, std::vector<double>& boxes) {
  const int ps = kind == KIND_AABB ? 2 * dim : dim + 1;
  boxes.resize((size_t)n * 2 * dim);
  for (uint64_t i = 0; i < n; ++i) {
    const double* p = prims + i * ps;
    double* bx = boxes.data() + i * 2 * dim;
    for (int k = 0; k < dim; ++k) {
      if (kind == KIND_AABB) { bx[k] = p[k]; bx[dim + k] = p[dim + k]; }
      else { bx[k] = p[k] - p[dim]; bx[dim + k] = p[k] + p[dim]; }
    }
    if (kind == KIND_SPHERE && !(p[dim] >= 0.0)) return fail(SQ_ERR_INVALID, "sphere radius must be >= 0");
  }
  return SQ_OK;
}

int sq_bvh64_create(sq_ctx* c, int dim, int kind, const double* prims, uint64_t n, sq_bvh64** out) {
  if (!c || !out || (n && !prims)) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (n >= 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "too many primitives (ids are u32)");
  int rc = dispatch64(dim, kind, [&]<int D, int KIND>() -> int { return SQ_OK; });
  if (rc) return rc;
  std::vector<double> boxes;
  if ((rc = prim_boxes64(dim, kind, prims, n, boxes))) return rc;
  return create64(c, dim, kind, prims, kind == KIND_AABB ? 2 * dim : dim + 1, boxes, n, out);
}

// Bvh::refit (bvh.rs:45-48) under the `f64` feature: the same number of
// primitives at new positions (host, caller order); the topology is kept,
// bounds refit in double (tree.rs:54-75), records uploaded again.
int sq_bvh64_refit(sq_bvh64* b, const double* prims) {
  if (!b || (b->n_prims && !prims)) return fail(SQ_ERR_INVALID, "NULL argument");
  if (b->kind != KIND_AABB && b->kind != KIND_SPHERE)
    return fail(SQ_ERR_UNSUPPORTED, "sq_bvh64_refit takes trees over AABB or SPHERE primitives (instance trees: sq_tlas64_refit)");
  std::vector<double> boxes;
  int rc;
  if ((rc = prim_boxes64(b->dim, b->kind, prims, b->n_prims, boxes))) return rc;
  refit_bounds64(b->tree, boxes);
  return upload64(b, prims);
}

// Tlas::build (tlas.rs:35-56) over box or ball instances in double: each world
// box is Isometry::transform_aabb of the local box, [-half, half] or
// [-r, r]^D (isometry.rs:84-97).  Balls store splat(r) as their half extents.
static int inst_records64(int dim, bool ball, uint64_t n, const double* rot, const double* trans, const double* halves,
                          std::vector<double>& recs, std::vector<double>& boxes);

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
  std::vector<double> boxes, recs;
  int rc = inst_records64(dim, ball, n, rot, trans, halves, recs, boxes);
  if (rc) return rc;
  sq_bvh64* b = nullptr;
  if ((rc = create64(c, dim, ball ? KIND_BALL64 : KIND_INST64, recs.data(), dim * dim + 4 * dim, boxes, n, &b))) return rc;
  b->inst_half.assign(halves, halves + n * (ball ? 1 : dim));
  *out = reinterpret_cast<sq_tlas64*>(b);
  return SQ_OK;
}

// Instance records [n][D*D + 4D] (R, t, half extents, world box) and world
// boxes [n][2D] of box or ball instances: each world box is
// Isometry::transform_aabb of the local box, [-half, half] or [-r, r]^D
// (isometry.rs:84-97).  Balls store splat(r) as their half extents.
static int inst_records64(int dim, bool ball, uint64_t n, const double* rot, const double* trans, const double* halves,
                          std::vector<double>& recs, std::vector<double>& boxes) {
  const int D = dim, ps = D * D + 4 * D;
  boxes.resize((size_t)n * 2 * D);
  recs.resize((size_t)n * ps);
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
  return SQ_OK;
}

// Tlas::refit (tlas.rs:49-54) under the `f64` feature: new transforms (and
// half extents or radii, or NULL to keep them), host; world boxes by
// transform_aabb in double, the top tree refit (tree.rs:54-75), records
// uploaded again.
int sq_tlas64_refit(sq_tlas64* tl, const double* rot, const double* trans, const double* halves) {
  sq_bvh64* b = reinterpret_cast<sq_bvh64*>(tl);
  if (!b || (b->n_prims && (!rot || !trans))) return fail(SQ_ERR_INVALID, "NULL argument");
  if (b->kind != KIND_INST64 && b->kind != KIND_BALL64) return fail(SQ_ERR_INVALID, "not an instance tree");
  const bool ball = b->kind == KIND_BALL64;
  const uint64_t n = b->n_prims;
  std::vector<double> boxes, recs;
  int rc = inst_records64(b->dim, ball, n, rot, trans, halves ? halves : b->inst_half.data(), recs, boxes);
  if (rc) return rc;   // (nothing changed)
  if (halves) b->inst_half.assign(halves, halves + n * (ball ? 1 : b->dim));
  refit_bounds64(b->tree, boxes);
  return upload64(b, recs.data());
}

// Tlas::needs_rebuild (tlas.rs:58-60): the top tree's interior surface sum
// now against factor times its value at the build.
int sq_tlas64_needs_rebuild(sq_tlas64* tl, double factor, int* out) {
  sq_bvh64* b = reinterpret_cast<sq_bvh64*>(tl);
  if (!b || !out) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = (b->baseline > 0.0 && surface_sum64(b->tree) > factor * b->baseline) ? 1 : 0;
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

// ---------------------------------------------------------------------------
// Consumer narrow phases in double: the f64 feature's QueryGeometry::test_ray
// (geometry.rs:33-43) for any provider, compiled at run time (hipRTC, as
// rtc.hip compiles the f32 ones) into the pooled f64 walk of walk64.h.  The
// consumer defines
//   __device__ bool sq_narrow64(const double* prim, const double* origin,
//                               const double* dir, double max_toi, double* toi);
// over its SQ_RTC_PF doubles per primitive; a hit counts in [0, max_toi], as the
// reference counts it.  Nearest queries walk the reference's order with the
// consumer's unguarded hits (the reference's answer, hits outside their own
// box included); under SQ_HITS_GUARDED a hit is kept only inside the slab
// interval of its world box and the walk is near-first.  All hits keep every
// hit in both modes, as the built-in f64 kinds' lists do.
// ---------------------------------------------------------------------------

namespace {

const char* kKernelSrc64 = R"SQRTC(
namespace sqd {
template <int D, bool GUARD, bool XC = false> struct ConsumerPol64 {
  static constexpr int QS = 2 * D;
  static constexpr bool CROSS = XC;   // all-hits lists of every crossing (test_ray_crossings)
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
    double tt;
    if (!sq_narrow64(p + 2 * D, r.o, r.d, limit, &tt)) return false;
    if (!(tt >= 0.0 && tt <= limit)) return false;
    if constexpr (GUARD) {
      double te, tx; int ax;
      if (!slab_exact<D>(r, p, p + D, limit, te, tx, ax)) return false;
      if (!(te <= tt && tt <= tx)) return false;
    }
    t = tt;
    return true;
  }
  // test_ray_crossings (geometry.rs:45-63): the source's sq_narrow64_crossings
  // when it defines one (SQ_NARROW64_CROSSINGS), each crossing kept under the
  // test_ray contract ([0, limit]; with GUARD, inside its box's slab
  // interval); else the trait's default, the one test_ray hit
  template <class Emit> __device__ void crossings(const double* p, double limit, Emit&& emit) const {
#if defined(SQ_NARROW64_CROSSINGS)
    auto out = [&](double tt) {
      if (!(tt >= 0.0 && tt <= limit)) return;
      if constexpr (GUARD) {
        double te, tx; int ax;
        if (!slab_exact<D>(r, p, p + D, limit, te, tx, ax) || !(te <= tt && tt <= tx)) return;
      }
      emit(tt);
    };
    sq_narrow64_crossings(p + 2 * D, r.o, r.d, limit, out);
#else
    double tt;
    if (narrow(p, limit, tt)) emit(tt);
#endif
  }
};
template <bool EXACT, bool ALL, bool XC = false>
__device__ __forceinline__ void consumer64(const View64& v, const double* rays, unsigned n, double max_toi, unsigned* counter,
                                           int* out_leaf, double* out_toi, SpillEntry* spill, unsigned lanes, unsigned* counts,
                                           const unsigned long long* offsets) {
  const Filter f{nullptr, ~0ull, -1};
  pool64_walk<SQ_RTC_D, KIND_CONSUMER64, false, false, ALL, EXACT, ConsumerPol64<SQ_RTC_D, !EXACT, XC>, 2 * SQ_RTC_D + SQ_RTC_PF>(
      v, rays, n, max_toi, nullptr, f, counter, out_leaf, out_toi, nullptr, nullptr, nullptr, spill, lanes, counts,
      reinterpret_cast<const uint64_t*>(offsets));
}
}  // namespace sqd

extern "C" __global__ __launch_bounds__(sqd::F64_BLOCK) void sq_consumer64_nearest(
    sqd::View64 v, const double* rays, unsigned n, double max_toi, unsigned* counter, int* out_leaf, double* out_toi,
    sqd::SpillEntry* spill, unsigned lanes) {
  sqd::consumer64<true, false>(v, rays, n, max_toi, counter, out_leaf, out_toi, spill, lanes, nullptr, nullptr);
}
extern "C" __global__ __launch_bounds__(sqd::F64_BLOCK) void sq_consumer64_nearest_guarded(
    sqd::View64 v, const double* rays, unsigned n, double max_toi, unsigned* counter, int* out_leaf, double* out_toi,
    sqd::SpillEntry* spill, unsigned lanes) {
  sqd::consumer64<false, false>(v, rays, n, max_toi, counter, out_leaf, out_toi, spill, lanes, nullptr, nullptr);
}
// all hits: counts[q] = hits (offsets null), or the hits written into the
// query's CSR range and sorted
extern "C" __global__ __launch_bounds__(sqd::F64_BLOCK) void sq_consumer64_all(
    sqd::View64 v, const double* rays, unsigned n, double max_toi, unsigned* counter, int* out_leaf, double* out_toi,
    sqd::SpillEntry* spill, unsigned lanes, unsigned* counts, const unsigned long long* offsets) {
  sqd::consumer64<true, true>(v, rays, n, max_toi, counter, out_leaf, out_toi, spill, lanes, counts, offsets);
}
// all crossings (Bvh::raycast_crossings): as sq_consumer64_all, every crossing of every primitive
extern "C" __global__ __launch_bounds__(sqd::F64_BLOCK) void sq_consumer64_crossings(
    sqd::View64 v, const double* rays, unsigned n, double max_toi, unsigned* counter, int* out_leaf, double* out_toi,
    sqd::SpillEntry* spill, unsigned lanes, unsigned* counts, const unsigned long long* offsets) {
  sqd::consumer64<true, true, true>(v, rays, n, max_toi, counter, out_leaf, out_toi, spill, lanes, counts, offsets);
}
)SQRTC";

}  // namespace

struct sq_narrow64 {
  sq_ctx* ctx = nullptr;
  int dim = 0;
  uint32_t pd = 0;
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;       // nearest hit
  hipFunction_t fn_g = nullptr;     // nearest hit, SQ_HITS_GUARDED
  hipFunction_t fn_all = nullptr;   // all hits (count, then fill)
  hipFunction_t fn_x = nullptr;     // all crossings (count, then fill)
  int per_cu[4] = {-1, -1, -1, -1}; // resident blocks per CU of fn, fn_g, fn_all, fn_x (queried at creation)
};

namespace {

// The tree was built for this narrow phase (same dimension, record size and
// context) and holds consumer records.
int check_pair64(const sq_narrow64* nw, const sq_bvh64* b) {
  if (b->kind != KIND_CONSUMER64 || b->dim != nw->dim || b->consumer_pd != nw->pd || b->ctx != nw->ctx)
    return fail(SQ_ERR_INVALID, "the tree was not built for this narrow phase");
  return SQ_OK;
}

// One pooled launch of module kernel fn (index k of sq_narrow64::per_cu) over
// n rays: a grid of the resident blocks, lanes refilled from the context's
// pool counter, ring stacks spilling (as pool_launch for the built-in kinds).
int consumer64_launch(sq_narrow64* nw, sq_bvh64* b, hipFunction_t fn, int k, const double* rays, uint64_t n,
                      double max_toi, int32_t* out_leaf, double* out_toi, uint32_t* counts, const uint64_t* offsets) {
  sq_ctx* c = b->ctx;
  const size_t lds = (size_t)F64_RING * F64_BLOCK * 8;
  const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)c->cus * std::max(1, nw->per_cu[k]),
                                                                 grid_for(n, F64_BLOCK)));
  uint32_t lanes = (uint32_t)(grid * F64_BLOCK);
  const uint64_t rows = b->view.stack > F64_RING ? b->view.stack - F64_RING : 0;
  if (int r1 = c->spill.ensure((size_t)rows * lanes * sizeof(SpillEntry) + 64)) return r1;
  View64 v = b->view;
  const double* rp = rays;
  uint32_t nn = (uint32_t)n;
  uint32_t* counter = c->pool_ctr.as<uint32_t>();
  SpillEntry* spill = c->spill.as<SpillEntry>();
  int32_t* lp = out_leaf;
  double* tp = out_toi;
  uint32_t* cp = counts;
  const uint64_t* op = offsets;
  void* args_nearest[] = {&v, &rp, &nn, &max_toi, &counter, &lp, &tp, &spill, &lanes};
  void* args_all[] = {&v, &rp, &nn, &max_toi, &counter, &lp, &tp, &spill, &lanes, &cp, &op};
  HIP_TRY(hipModuleLaunchKernel(fn, (unsigned)grid, 1, 1, F64_BLOCK, 1, 1, (unsigned)lds, c->stream,
                                k >= 2 ? args_all : args_nearest, nullptr));   // 2, 3: the list kernels (11 arguments)
  return SQ_OK;
}

}  // namespace

extern "C" {

int sq_narrow64_create(sq_ctx* ctx, int dim, const char* source, uint32_t prim_doubles, sq_narrow64** out) {
  if (!ctx || !source || !out) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (dim < 2 || dim > 6) return fail(SQ_ERR_UNSUPPORTED, "consumer narrow phases take dim 2..6");
  if (prim_doubles > 128) return fail(SQ_ERR_INVALID, "at most 128 doubles per primitive");
  auto nw = std::make_unique<sq_narrow64>();
  nw->ctx = ctx;
  nw->dim = dim;
  nw->pd = prim_doubles;
  if (int rc = rtc_load_module(ctx, "walk64.h", source, kKernelSrc64, dim, prim_doubles, &nw->mod)) return rc;
  if (hipModuleGetFunction(&nw->fn, nw->mod, "sq_consumer64_nearest") != hipSuccess ||
      hipModuleGetFunction(&nw->fn_g, nw->mod, "sq_consumer64_nearest_guarded") != hipSuccess ||
      hipModuleGetFunction(&nw->fn_all, nw->mod, "sq_consumer64_all") != hipSuccess ||
      hipModuleGetFunction(&nw->fn_x, nw->mod, "sq_consumer64_crossings") != hipSuccess) {
    (void)hipModuleUnload(nw->mod);
    return fail(SQ_ERR_HIP, "consumer kernel not found in the compiled module");
  }
  const size_t lds = (size_t)F64_RING * F64_BLOCK * 8;
  const hipFunction_t fns[4] = {nw->fn, nw->fn_g, nw->fn_all, nw->fn_x};
  for (int k = 0; k < 4; ++k) {
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nw->per_cu[k], fns[k], F64_BLOCK, lds) != hipSuccess) {
      (void)hipModuleUnload(nw->mod);
      return fail(SQ_ERR_HIP, "occupancy query of the consumer kernels failed");
    }
  }
  ctx_retain(ctx);
  *out = nw.release();
  return SQ_OK;
}

int sq_narrow64_destroy(sq_narrow64* nw) {
  if (!nw) return SQ_OK;
  sq_ctx* c = nw->ctx;
  (void)hipStreamSynchronize(c->stream);
  if (nw->mod) (void)hipModuleUnload(nw->mod);
  delete nw;
  ctx_release(c);
  return SQ_OK;
}

// The f64 tree over n consumer primitives: world boxes [n][2D] (host, the
// median split's input, QueryGeometry::world_aabb) and the consumer's doubles
// [n][prim_doubles] (host); each leaf-ordered device record is the box, then
// the doubles.
int sq_bvh64_build_consumer(sq_narrow64* nw, const double* boxes, const double* prims, uint64_t n, sq_bvh64** out) {
  if (!nw || !out || (n && (!boxes || (nw->pd && !prims)))) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (n >= 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "too many primitives (ids are u32)");
  const int D = nw->dim;
  const int rs = 2 * D + (int)nw->pd;
  std::vector<double> bx(boxes, boxes + (size_t)n * 2 * D);
  std::vector<double> recs((size_t)n * rs);
  for (uint64_t i = 0; i < n; ++i) {
    for (int k = 0; k < D; ++k)
      if (!(bx[i * 2 * D + k] <= bx[i * 2 * D + D + k])) return fail(SQ_ERR_INVALID, "a world box has min > max (or NaN)");
    std::memcpy(recs.data() + i * rs, boxes + i * 2 * D, 2 * D * sizeof(double));
    if (nw->pd) std::memcpy(recs.data() + i * rs + 2 * D, prims + i * nw->pd, nw->pd * sizeof(double));
  }
  sq_bvh64* b = nullptr;
  if (int rc = create64(nw->ctx, D, KIND_CONSUMER64, recs.data(), rs, bx, n, &b)) return rc;
  b->consumer_pd = nw->pd;
  *out = b;
  return SQ_OK;
}

int sq_raycast64_nearest_consumer(sq_narrow64* nw, sq_bvh64* b, const double* rays, uint64_t n, double max_toi,
                                  int32_t* out_leaf, double* out_toi) {
  if (!nw || !b || (n && (!rays || !out_leaf || !out_toi))) return fail(SQ_ERR_INVALID, "NULL argument");
  if (int rc = check_pair64(nw, b)) return rc;
  if (!n) return SQ_OK;
  if (n >= (1ull << 31)) return fail(SQ_ERR_INVALID, "more than 2^31 queries in one call");
  const bool guarded = b->ctx->hit_mode == SQ_HITS_GUARDED;
  return consumer64_launch(nw, b, guarded ? nw->fn_g : nw->fn, guarded ? 1 : 0, rays, n, max_toi, out_leaf, out_toi,
                           nullptr, nullptr);
}

// all hits (kernel 2) or all crossings (kernel 3): count pass, scan, fill pass
static int consumer64_lists(sq_narrow64* nw, sq_bvh64* b, int which, const double* rays, uint64_t n, double max_toi,
                            uint64_t* offsets, int32_t* out_leaf, double* out_toi, uint64_t capacity, uint64_t* total) {
  if (!nw || !b || !offsets || !total || (n && !rays)) return fail(SQ_ERR_INVALID, "NULL argument");
  const hipFunction_t fk = which == 3 ? nw->fn_x : nw->fn_all;
  if (int rc = check_pair64(nw, b)) return rc;
  sq_ctx* c = b->ctx;
  *total = 0;
  if (!n) {
    HIP_TRY(hipMemsetAsync(offsets, 0, 8, c->stream));
    return SQ_OK;
  }
  if (n >= (1ull << 31)) return fail(SQ_ERR_INVALID, "more than 2^31 queries in one call");
  int rc;
  if ((rc = c->counts.ensure(n * 4))) return rc;
  uint32_t* counts = c->counts.as<uint32_t>();
  if ((rc = consumer64_launch(nw, b, fk, which, rays, n, max_toi, nullptr, nullptr, counts, nullptr))) return rc;
  std::vector<std::unique_ptr<DevBuf>> temps;
  if ((rc = scan_counts(c, counts, n, offsets, temps))) return rc;
  HIP_TRY(hipMemcpyAsync(total, offsets + n, 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const uint64_t tot = *total;
  if (tot > capacity)
    return fail(SQ_ERR_CAPACITY, "result needs " + std::to_string(tot) + " entries, capacity is " + std::to_string(capacity));
  if (!tot) return SQ_OK;
  if (!out_leaf || !out_toi) return fail(SQ_ERR_INVALID, "output arrays are NULL");
  return consumer64_launch(nw, b, fk, which, rays, n, max_toi, out_leaf, out_toi, counts, offsets);
}

int sq_raycast64_all_consumer(sq_narrow64* nw, sq_bvh64* b, const double* rays, uint64_t n, double max_toi,
                              uint64_t* offsets, int32_t* out_leaf, double* out_toi, uint64_t capacity, uint64_t* total) {
  return consumer64_lists(nw, b, 2, rays, n, max_toi, offsets, out_leaf, out_toi, capacity, total);
}

int sq_raycast64_crossings_consumer(sq_narrow64* nw, sq_bvh64* b, const double* rays, uint64_t n, double max_toi,
                                    uint64_t* offsets, int32_t* out_leaf, double* out_toi, uint64_t capacity,
                                    uint64_t* total) {
  return consumer64_lists(nw, b, 3, rays, n, max_toi, offsets, out_leaf, out_toi, capacity, total);
}

}  // extern "C"

