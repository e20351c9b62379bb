// sq_oracle64.cpp -- CPU oracle of the reference's `f64` feature.
//
// TEST INFRASTRUCTURE ONLY (the same rules as sq_oracle.cpp: only `tests/` load
// it, as the checker).  Under the feature the reference's Scalar is f64
// (sq/src/maths/point.rs:14-18), so every operation below is the double
// restatement of the f32 one, in the reference's order, with FP contraction
// disabled.  It restates:
//   * Aabb::ray_intersect          sq/src/maths/aabb.rs:118-148 (f64::EPSILON)
//   * Aabb::intersects             sq/src/maths/aabb.rs:101-108
//   * the naive scans              sq/src/query/naive.rs:34-62 (nearest),
//                                  :65-109 (all hits, sorted by hit_ordering),
//                                  :131-181 (shape casts, nearest and all),
//                                  :184-205 (overlap)
//   * ShapeCast broadphase         sq/src/query/shapecast.rs:70-73 (box grown by the probe)
//   * box instances                two_level.rs:70-113 (ray_local_box), the naive instance scan
//                                  of two_level.rs:126-167, Isometry inverse and transform_aabb
//                                  (isometry.rs:20-124), kept inside the world box's slab interval
//   * ball instances               shape_overlap.rs:288-345 (BallInstances: ray_ball on the local
//                                  ray :120-138, the cast against the radius sum :326-337,
//                                  point_to_aabb overlap :59-72,339-341)
//   * hit_ordering                 sq/src/query/hit.rs:77-84 (f64::total_cmp, then leaf)
//   * tree::build and TreeStats    sq/src/accel/tree.rs:39-49,81-163
//   * Bvh::raycast_nearest         sq/src/accel/bvh.rs:51-124 (push a, push b, LIFO, shrinking
//                                  limit) with the QueryStats counters of stats.rs:39-50,
//                                  over raycast_nearest_batch_parallel's contiguous chunks
//                                  (bvh.rs:160-193): the timed CPU baseline
//   * the Bvh walks of every query  bvh.rs:196-259 (all hits), :331-414 (shape casts),
//                                  :418-456 (overlap), and the Tlas walks tlas.rs:104-154
//                                  (nearest), :158-215 (all hits), :221-304 (crossings: one per
//                                  instance by test_ray_crossings_local's default), :306-391
//                                  (shape casts), :393-432 (overlap) over box and ball
//                                  instances: the expected query results
// over the built-in BOX and SPHERE providers, and over a consumer provider
// (QueryGeometry::test_ray of geometry.rs:33-43 in double, a C function built
// from the consumer's source: orc64_consumer_query).  A `kind` with FLAG_GUARDED set
// restates the engine's SQ_HITS_GUARDED mode instead (a sphere or instance hit
// kept only inside the slab interval of its own box, so that BVH == naive scan,
// the reference's stated contract, sq/src/lib.rs:15-17, holds exactly); with
// FLAG_BVH the queries run the reference's Bvh walks, else the naive scans.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace orc64 {

static const double F64_EPSILON = 2.220446049250313e-16;  // f64::EPSILON
static const uint32_t LEAF_SIZE = 4;                      // sq/src/accel/tree.rs:12
// Flags or-ed into the `kind` argument of the C entry points.
static const int FLAG_GUARDED = 0x100;   // the engine's SQ_HITS_GUARDED leaf tests
static const int FLAG_BVH = 0x200;       // the reference's Bvh / Tlas walks instead of the naive scans

struct Ray { const double* o; const double* d; };

// aabb.rs:118-148
template <int D>
static bool ray_intersect(const double* mn, const double* mx, const Ray& r, double max_toi, double* tmin, double* tmax,
                          int* axis_out) {
  double t_min = 0.0, t_max = max_toi;
  int axis = -1;
  for (int i = 0; i < D; ++i) {
    const double o = r.o[i], d = r.d[i];
    if (std::fabs(d) < F64_EPSILON) {
      if (o < mn[i] || o > mx[i]) return false;
    } else {
      const double inv = 1.0 / d;
      double t1 = (mn[i] - o) * inv;
      double t2 = (mx[i] - o) * inv;
      if (t1 > t2) std::swap(t1, t2);
      if (t1 > t_min) { t_min = t1; axis = i; }
      if (t2 < t_max) t_max = t2;
      if (t_min > t_max) return false;
    }
  }
  *tmin = t_min; *tmax = t_max; *axis_out = axis;
  return true;
}

// point.rs:60-66: accumulate from 0.0, left to right.
template <int D> static double dot(const double* a, const double* b) {
  double acc = 0.0;
  for (int i = 0; i < D; ++i) acc = acc + a[i] * b[i];
  return acc;
}

// A consumer's test_ray in double (geometry.rs:33-43): nonzero with *toi on a
// hit of the primitive's doubles.
typedef int (*ConsumerNarrow64)(const double* prim, const double* origin, const double* dir, double max_toi, double* toi);
// A consumer's test_ray_crossings in double (geometry.rs:45-63): calls
// emit(ctx, toi) once per crossing of the ray with the primitive.
typedef void (*ConsumerCrossings64)(const double* prim, const double* origin, const double* dir, double max_toi,
                                    void (*emit)(void*, double), void* ctx);

template <int D> struct Geom {
  int kind;   // 0 BOX [min D][max D], 1 SPHERE [c D][r], 2 CONSUMER [world box 2D][consumer_pd doubles]
  uint32_t n;
  const double* prims;
  const uint64_t* layers;
  bool guarded = false;   // sphere and consumer hits (and contacts) kept inside their box's slab interval
  ConsumerNarrow64 consumer = nullptr;
  ConsumerCrossings64 consumer_x = nullptr;   // set: all-hits lists report every crossing (raycast_crossings)
  int consumer_pd = 0;
  int stride() const { return kind == 0 ? 2 * D : kind == 1 ? D + 1 : 2 * D + consumer_pd; }
  const double* p(size_t leaf) const { return prims + leaf * stride(); }
  // QueryGeometry::world_aabb: the box (BOX, and the consumer's given box), or
  // Aabb::new(c - splat(r), c + splat(r))
  void world_aabb(size_t leaf, double* mn, double* mx) const {
    const double* q = p(leaf);
    for (int i = 0; i < D; ++i) {
      if (kind == 0 || kind == 2) { mn[i] = q[i]; mx[i] = q[D + i]; }
      else { mn[i] = q[i] - q[D]; mx[i] = q[i] + q[D]; }
    }
  }
  // The leaf's hits for an all-hits list: every crossing the consumer reports
  // (bvh.rs:264-328 calling test_ray_crossings), each kept under the test_ray
  // contract and, when `guarded`, inside its box's slab interval; else the
  // test_ray hit.
  template <class F> void list_hits(size_t leaf, const Ray& r, double max_toi, F&& out) const {
    if (kind == 2 && consumer_x) {
      std::vector<double> ts;
      consumer_x(p(leaf) + 2 * D, r.o, r.d, max_toi, [](void* c, double t) { static_cast<std::vector<double>*>(c)->push_back(t); },
                 &ts);
      double mn[D], mx[D], te, tx; int ax;
      world_aabb(leaf, mn, mx);
      for (double t : ts) {
        if (!(t >= 0.0 && t <= max_toi)) continue;
        if (guarded && (!ray_intersect<D>(mn, mx, r, max_toi, &te, &tx, &ax) || !(te <= t && t <= tx))) continue;
        out(t);
      }
      return;
    }
    double t, nrm[D];
    if (test_ray(leaf, r, max_toi, &t, nrm)) out(t);
  }
  bool accepts(size_t leaf, uint64_t mask, int64_t exclude) const {
    if (exclude >= 0 && (int64_t)leaf == exclude) return false;
    const uint64_t lay = layers ? layers[leaf] : ~0ull;
    return (lay & mask) != 0;
  }
  // test_ray: BOX answers the slab entry with the entry face normal; SPHERE the
  // ball (entry root, the exit root from inside), guarded by its own box when
  // `guarded`.
  bool test_ray(size_t leaf, const Ray& r, double max_toi, double* t_out, double* normal) const {
    const double* q = p(leaf);
    double mn[D], mx[D];
    world_aabb(leaf, mn, mx);
    double te, tx; int axis;
    if (kind == 2) {
      // the consumer's hit, counted in [0, max_toi] (the test_ray contract, geometry.rs:37-39)
      double t;
      if (!consumer(q + 2 * D, r.o, r.d, max_toi, &t)) return false;
      if (!(t >= 0.0 && t <= max_toi)) return false;
      if (guarded && !ray_intersect<D>(mn, mx, r, max_toi, &te, &tx, &axis)) return false;
      if (guarded && !(te <= t && t <= tx)) return false;
      *t_out = t;
      for (int i = 0; i < D; ++i) normal[i] = 0.0;
      return true;
    }
    if (kind == 0) {
      if (!ray_intersect<D>(mn, mx, r, max_toi, &te, &tx, &axis)) return false;
      *t_out = te;
      for (int i = 0; i < D; ++i) normal[i] = 0.0;
      if (axis >= 0) normal[axis] = r.d[axis] > 0.0 ? -1.0 : 1.0;
      return true;
    }
    double oc[D];
    for (int i = 0; i < D; ++i) oc[i] = r.o[i] - q[i];
    const double rad = q[D];
    const double b = dot<D>(oc, r.d);
    const double cc = dot<D>(oc, oc) - rad * rad;
    const double disc = b * b - cc;
    if (disc < 0.0) return false;
    const double sq = std::sqrt(disc);
    double t = -b - sq;
    if (t < 0.0) t = -b + sq;
    if (t < 0.0 || t > max_toi) return false;
    if (guarded && !ray_intersect<D>(mn, mx, r, max_toi, &te, &tx, &axis)) return false;
    if (guarded && !(te <= t && t <= tx)) return false;
    *t_out = t;
    // point.rs:82-89 normalize_or_zero(at(t) - c)
    double v[D];
    for (int i = 0; i < D; ++i) v[i] = (r.o[i] + r.d[i] * t) - q[i];
    const double len = std::sqrt(dot<D>(v, v));
    for (int i = 0; i < D; ++i) normal[i] = len > F64_EPSILON ? v[i] * (1.0 / len) : 0.0;
    return true;
  }
  // shapecast.rs:70-73: the slab against `mn, mx` grown by the probe box
  bool broadphase(const double* mn, const double* mx, const Ray& r, const double* pmin, const double* pmax, double max_toi,
                  double* te, double* tx) const {
    double gmn[D], gmx[D];
    for (int i = 0; i < D; ++i) { gmn[i] = mn[i] - pmax[i]; gmx[i] = mx[i] - pmin[i]; }
    int ax;
    return ray_intersect<D>(gmn, gmx, r, max_toi, te, tx, &ax);
  }
  // test_shape_cast: BOX the grown slab's entry; SPHERE the ball of radius
  // r + probe.max[0] (shape_overlap.rs:104-118), guarded by the grown box when
  // `guarded`
  bool test_cast(size_t leaf, const Ray& r, const double* pmin, const double* pmax, double max_toi, double* t_out) const {
    const double* q = p(leaf);
    double mn[D], mx[D], te, tx;
    world_aabb(leaf, mn, mx);
    if (kind == 0) {
      if (!broadphase(mn, mx, r, pmin, pmax, max_toi, &te, &tx)) return false;
      *t_out = te;
      return true;
    }
    double oc[D];
    for (int i = 0; i < D; ++i) oc[i] = r.o[i] - q[i];
    const double rad = q[D] + pmax[0];
    const double b = dot<D>(oc, r.d);
    const double cc = dot<D>(oc, oc) - rad * rad;
    const double disc = b * b - cc;
    if (disc < 0.0) return false;
    const double sq = std::sqrt(disc);
    double t = -b - sq;
    if (t < 0.0) t = -b + sq;
    if (t < 0.0 || t > max_toi) return false;
    if (guarded && !broadphase(mn, mx, r, pmin, pmax, max_toi, &te, &tx)) return false;
    if (guarded && !(te <= t && t <= tx)) return false;
    *t_out = t;
    return true;
  }
  // test_overlap: the box test; the ball's nearest region point within r
  bool test_overlap(size_t leaf, const double* rmn, const double* rmx) const {
    const double* q = p(leaf);
    if (kind == 0) {
      for (int i = 0; i < D; ++i) if (q[i] > rmx[i] || q[D + i] < rmn[i]) return false;
      return true;
    }
    double d2 = 0.0;
    for (int i = 0; i < D; ++i) {
      double v;
      if (q[i] < rmn[i]) v = rmn[i] - q[i];
      else if (q[i] > rmx[i]) v = q[i] - rmx[i];
      else v = 0.0;
      d2 = d2 + v * v;
    }
    return std::sqrt(d2) <= q[D];
  }
};

static int64_t total_key(double f) {
  int64_t i;
  std::memcpy(&i, &f, 8);
  return i ^ (int64_t)(((uint64_t)(i >> 63)) >> 1);
}
// hit.rs:77-84
static bool hit_less(double ta, int64_t la, double tb, int64_t lb) {
  const int64_t a = total_key(ta), b = total_key(tb);
  return a < b || (a == b && la < lb);
}

// naive.rs:34-62
template <int D>
static bool naive_nearest(const Geom<D>& g, const Ray& r, double max_toi, uint64_t mask, int64_t exclude, int64_t* leaf_out,
                          double* t_out, double* normal_out) {
  bool have = false;
  double limit = max_toi, best_t = 0.0;
  int64_t best_leaf = -1;
  double nrm[D];
  for (size_t leaf = 0; leaf < g.n; ++leaf) {
    if (!g.accepts(leaf, mask, exclude)) continue;
    double mn[D], mx[D], te, tx; int ax;
    g.world_aabb(leaf, mn, mx);
    if (!ray_intersect<D>(mn, mx, r, limit, &te, &tx, &ax)) continue;   // the cheap reject
    double t;
    if (!g.test_ray(leaf, r, limit, &t, nrm)) continue;
    if (have && !hit_less(t, (int64_t)leaf, best_t, best_leaf)) continue;
    have = true; best_t = t; best_leaf = (int64_t)leaf; limit = t;
    for (int i = 0; i < D; ++i) normal_out[i] = nrm[i];
  }
  *leaf_out = best_leaf; *t_out = best_t;
  return have;
}

// naive.rs:65-109 (sorted nearest-first, then by leaf)
template <int D>
static void naive_all(const Geom<D>& g, const Ray& r, double max_toi, uint64_t mask, int64_t exclude,
                      std::vector<std::pair<double, int64_t>>& out) {
  out.clear();
  double nrm[D];
  for (size_t leaf = 0; leaf < g.n; ++leaf) {
    if (!g.accepts(leaf, mask, exclude)) continue;
    double mn[D], mx[D], te, tx; int ax;
    g.world_aabb(leaf, mn, mx);
    if (!ray_intersect<D>(mn, mx, r, max_toi, &te, &tx, &ax)) continue;
    g.list_hits(leaf, r, max_toi, [&](double t) { out.push_back({t, (int64_t)leaf}); });
  }
  (void)nrm;
  std::stable_sort(out.begin(), out.end(),
                   [](const std::pair<double, int64_t>& a, const std::pair<double, int64_t>& b) {
                     return hit_less(a.first, a.second, b.first, b.second);
                   });
}

// naive.rs:131-158 (nearest) and :161-181 (all, sorted by hit_ordering)
template <int D>
static bool naive_cast_nearest(const Geom<D>& g, const Ray& r, const double* pmin, const double* pmax, double max_toi,
                               uint64_t mask, int64_t exclude, int64_t* leaf_out, double* t_out) {
  bool have = false;
  double limit = max_toi, best_t = 0.0;
  int64_t best_leaf = -1;
  for (size_t leaf = 0; leaf < g.n; ++leaf) {
    if (!g.accepts(leaf, mask, exclude)) continue;
    double mn[D], mx[D], te, tx;
    g.world_aabb(leaf, mn, mx);
    if (!g.broadphase(mn, mx, r, pmin, pmax, limit, &te, &tx)) continue;
    double t;
    if (!g.test_cast(leaf, r, pmin, pmax, limit, &t)) continue;
    if (have && !hit_less(t, (int64_t)leaf, best_t, best_leaf)) continue;
    have = true; best_t = t; best_leaf = (int64_t)leaf; limit = t;
  }
  *leaf_out = best_leaf; *t_out = best_t;
  return have;
}

template <int D>
static void naive_cast_all(const Geom<D>& g, const Ray& r, const double* pmin, const double* pmax, double max_toi,
                           uint64_t mask, int64_t exclude, std::vector<std::pair<double, int64_t>>& out) {
  out.clear();
  for (size_t leaf = 0; leaf < g.n; ++leaf) {
    if (!g.accepts(leaf, mask, exclude)) continue;
    double mn[D], mx[D], te, tx;
    g.world_aabb(leaf, mn, mx);
    if (!g.broadphase(mn, mx, r, pmin, pmax, max_toi, &te, &tx)) continue;
    double t;
    if (g.test_cast(leaf, r, pmin, pmax, max_toi, &t)) out.push_back({t, (int64_t)leaf});
  }
  std::stable_sort(out.begin(), out.end(),
                   [](const std::pair<double, int64_t>& a, const std::pair<double, int64_t>& b) {
                     return hit_less(a.first, a.second, b.first, b.second);
                   });
}

// naive.rs:184-205
template <int D>
static void naive_overlap(const Geom<D>& g, const double* rmn, const double* rmx, uint64_t mask, int64_t exclude,
                          std::vector<int64_t>& out) {
  out.clear();
  for (size_t leaf = 0; leaf < g.n; ++leaf) {
    if (!g.accepts(leaf, mask, exclude)) continue;
    double mn[D], mx[D];
    g.world_aabb(leaf, mn, mx);
    bool inter = true;
    for (int i = 0; i < D; ++i) if (mn[i] > rmx[i] || mx[i] < rmn[i]) inter = false;
    if (!inter) continue;
    if (g.test_overlap(leaf, rmn, rmx)) out.push_back((int64_t)leaf);
  }
}

// tree.rs:112-163 in f64: union fold from EMPTY, leaf at <= 4, centroid bounds,
// longest axis (first max wins), coincidence stop at f64::EPSILON, stable sort
// by total_cmp of the centroid, split at count / 2.
struct Node { std::vector<double> b; uint32_t a, c, pc; };

template <int D>
static uint32_t build(const std::vector<double>& boxes, std::vector<uint32_t>& idx, uint32_t start, uint32_t end,
                      std::vector<Node>& nodes) {
  Node nd;
  nd.b.assign(2 * D, 0.0);
  for (int i = 0; i < D; ++i) { nd.b[i] = INFINITY; nd.b[D + i] = -INFINITY; }
  for (uint32_t k = start; k < end; ++k) {
    const double* b = boxes.data() + (size_t)idx[k] * 2 * D;
    for (int i = 0; i < D; ++i) { nd.b[i] = std::fmin(nd.b[i], b[i]); nd.b[D + i] = std::fmax(nd.b[D + i], b[D + i]); }
  }
  const uint32_t count = end - start;
  nd.a = start; nd.c = 0; nd.pc = count;
  const uint32_t me = (uint32_t)nodes.size();
  nodes.push_back(nd);
  if (count <= LEAF_SIZE) return me;
  double cmin[D], cmax[D];
  for (int i = 0; i < D; ++i) { cmin[i] = INFINITY; cmax[i] = -INFINITY; }
  auto centre = [&](uint32_t prim, int i) {
    const double* b = boxes.data() + (size_t)prim * 2 * D;
    return (b[i] + b[D + i]) * 0.5;
  };
  for (uint32_t k = start; k < end; ++k)
    for (int i = 0; i < D; ++i) { const double c = centre(idx[k], i); cmin[i] = std::fmin(cmin[i], c); cmax[i] = std::fmax(cmax[i], c); }
  int axis = 0;
  for (int i = 1; i < D; ++i) if (cmax[i] - cmin[i] > cmax[axis] - cmin[axis]) axis = i;
  if (cmax[axis] - cmin[axis] <= F64_EPSILON) return me;
  std::stable_sort(idx.begin() + start, idx.begin() + end,
                   [&](uint32_t x, uint32_t y) { return total_key(centre(x, axis)) < total_key(centre(y, axis)); });
  const uint32_t mid = start + count / 2;
  const uint32_t a = build<D>(boxes, idx, start, mid, nodes);
  const uint32_t c = build<D>(boxes, idx, mid, end, nodes);
  nodes[me].a = a; nodes[me].c = c; nodes[me].pc = 0;
  return me;
}

template <int D>
static void tree_stats(const double* prims, int kind, uint32_t n, uint64_t* out5, double* quality) {
  Geom<D> g{kind, n, prims, nullptr};
  std::vector<double> boxes((size_t)n * 2 * D);
  for (uint32_t i = 0; i < n; ++i) g.world_aabb(i, boxes.data() + (size_t)i * 2 * D, boxes.data() + (size_t)i * 2 * D + D);
  std::vector<uint32_t> idx(n);
  for (uint32_t i = 0; i < n; ++i) idx[i] = i;
  std::vector<Node> nodes;
  if (n) build<D>(boxes, idx, 0, n, nodes);
  // tree.rs:81-110
  uint64_t leaves = 0, pr = 0;
  std::vector<uint64_t> depth(nodes.size(), 0);
  double q = 0.0;
  for (size_t i = nodes.size(); i-- > 0;) {
    if (nodes[i].pc > 0) { ++leaves; pr += nodes[i].pc; depth[i] = 1; }
    else {
      depth[i] = 1 + std::max(depth[nodes[i].a], depth[nodes[i].c]);
    }
  }
  for (size_t i = 0; i < nodes.size(); ++i) {   // interior half surfaces in node order (aabb.rs:84-97)
    if (nodes[i].pc > 0) continue;
    double sum = 0.0;
    for (int a = 0; a < D; ++a) {
      double prod = 1.0;
      for (int k = 0; k < D; ++k) if (k != a) prod = prod * (nodes[i].b[D + k] - nodes[i].b[k]);
      sum = sum + prod;
    }
    q = q + sum;
  }
  out5[0] = nodes.size(); out5[1] = leaves; out5[2] = pr; out5[3] = nodes.empty() ? 0 : depth[0];
  out5[4] = nodes.size() * (uint64_t)(16 * D + 16) + (uint64_t)n * 4;
  *quality = q;
}

// A built f64 tree over one geometry: tree.rs:112-163 order, plus prim_indices.
template <int D> struct Tree64 {
  std::vector<Node> nodes;
  std::vector<uint32_t> idx;
  Tree64(const Geom<D>& g) {
    std::vector<double> boxes((size_t)g.n * 2 * D);
    for (uint32_t i = 0; i < g.n; ++i) g.world_aabb(i, boxes.data() + (size_t)i * 2 * D, boxes.data() + (size_t)i * 2 * D + D);
    init(boxes, g.n);
  }
  // over leaf boxes [n][min D, max D] (Tlas::build, tlas.rs:35-45: the world boxes)
  Tree64(const std::vector<double>& boxes, uint32_t n) { init(boxes, n); }
  void init(const std::vector<double>& boxes, uint32_t n) {
    idx.resize(n);
    for (uint32_t i = 0; i < n; ++i) idx[i] = i;
    if (n) build<D>(boxes, idx, 0, n, nodes);
  }
  // tree::refit (tree.rs:54-75) over fresh leaf boxes [n][min D, max D]: one
  // reverse pass over the nodes, a leaf folding its primitives' boxes from
  // EMPTY in prim_indices order, an interior node the union of its children
  // (Aabb::union, aabb.rs:38-43: componentwise min and max).
  void refit(const std::vector<double>& boxes) {
    for (size_t ni = nodes.size(); ni-- > 0;) {
      Node& nd = nodes[ni];
      if (nd.pc > 0) {
        for (int i = 0; i < D; ++i) { nd.b[i] = INFINITY; nd.b[D + i] = -INFINITY; }
        for (uint32_t k = nd.a; k < nd.a + nd.pc; ++k) {
          const double* b = boxes.data() + (size_t)idx[k] * 2 * D;
          for (int i = 0; i < D; ++i) { nd.b[i] = std::fmin(nd.b[i], b[i]); nd.b[D + i] = std::fmax(nd.b[D + i], b[D + i]); }
        }
      } else {
        const Node& a = nodes[nd.a];
        const Node& c = nodes[nd.c];
        for (int i = 0; i < D; ++i) { nd.b[i] = std::fmin(a.b[i], c.b[i]); nd.b[D + i] = std::fmax(a.b[D + i], c.b[D + i]); }
      }
    }
  }
  // tree::interior_surface_sum (tree.rs:104-110): half surfaces of the
  // interior nodes summed in node order.
  double quality() const {
    double q = 0.0;
    for (const Node& nd : nodes) {
      if (nd.pc > 0) continue;
      double sum = 0.0;
      for (int a = 0; a < D; ++a) {
        double prod = 1.0;
        for (int k = 0; k < D; ++k) if (k != a) prod = prod * (nd.b[D + k] - nd.b[k]);
        sum = sum + prod;
      }
      q = q + sum;
    }
    return q;
  }
  // The reference's walk (bvh.rs:73-124 and its siblings): pop a node, test it
  // with pass(min, max) (which reads the caller's current limit), then either
  // hand each primitive of a leaf to leaf(id) in prim_indices order, or push
  // child a, then child b.
  template <class Pass, class LeafFn> void walk(Pass&& pass, LeafFn&& leaf) const {
    if (nodes.empty()) return;
    std::vector<uint32_t> stack(1, 0u);
    while (!stack.empty()) {
      const uint32_t ni = stack.back();
      stack.pop_back();
      const Node& nd = nodes[ni];
      if (!pass(nd.b.data(), nd.b.data() + D)) continue;
      if (nd.pc > 0) {
        for (uint32_t k = nd.a; k < nd.a + nd.pc; ++k) leaf((size_t)idx[k]);
      } else {
        stack.push_back(nd.a);
        stack.push_back(nd.c);
      }
    }
  }
  // bvh.rs:51-124; counts nodes tested and primitives tested (stats.rs:39-50)
  bool nearest(const Geom<D>& g, const Ray& r, double max_toi, int64_t* leaf_out, double* t_out, uint64_t* n_nodes,
               uint64_t* n_prims) const {
    if (nodes.empty()) return false;
    bool have = false;
    double limit = max_toi, best_t = 0.0;
    int64_t best_leaf = -1;
    double nrm[D];
    std::vector<uint32_t> stack(1, 0u);
    while (!stack.empty()) {
      const uint32_t ni = stack.back();
      stack.pop_back();
      const Node& nd = nodes[ni];
      ++*n_nodes;
      double te, tx; int ax;
      if (!ray_intersect<D>(nd.b.data(), nd.b.data() + D, r, limit, &te, &tx, &ax) || !(te <= limit)) continue;
      if (nd.pc > 0) {
        for (uint32_t k = nd.a; k < nd.a + nd.pc; ++k) {
          const size_t leaf = idx[k];
          ++*n_prims;
          double t;
          if (!g.test_ray(leaf, r, limit, &t, nrm)) continue;
          if (have && !hit_less(t, (int64_t)leaf, best_t, best_leaf)) continue;
          have = true; best_t = t; best_leaf = (int64_t)leaf; limit = t;
        }
      } else {
        stack.push_back(nd.a);
        stack.push_back(nd.c);
      }
    }
    *leaf_out = best_leaf; *t_out = best_t;
    return have;
  }
};

template <class F> static void parallel_for(uint64_t n, int threads, F&& f) {
  if (threads <= 1 || n < 256) { for (uint64_t i = 0; i < n; ++i) f(i); return; }
  std::vector<std::thread> th;
  const uint64_t chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    const uint64_t a = t * chunk, b = std::min<uint64_t>(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([&, a, b] { for (uint64_t i = a; i < b; ++i) f(i); });
  }
  for (auto& t : th) t.join();
}

// The geometry of a `kind` argument with its flags.
template <int D> static Geom<D> geom_of(int kindf, uint32_t n, const double* prims, const uint64_t* layers) {
  Geom<D> g{kindf & 0xFF, n, prims, layers};
  g.guarded = (kindf & FLAG_GUARDED) != 0;
  return g;
}

// Bvh::raycast_nearest (bvh.rs:73-124) with the normal of the best hit.
template <int D>
static bool bvh_nearest(const Tree64<D>& tree, const Geom<D>& g, const Ray& r, double max_toi, uint64_t mask, int64_t exclude,
                        int64_t* leaf_out, double* t_out, double* normal_out) {
  bool have = false;
  double limit = max_toi, best_t = 0.0;
  int64_t best_leaf = -1;
  tree.walk([&](const double* mn, const double* mx) { double te, tx; int ax; return ray_intersect<D>(mn, mx, r, limit, &te, &tx, &ax); },
            [&](size_t leaf) {
              if (!g.accepts(leaf, mask, exclude)) return;
              double t, nrm[D];
              if (!g.test_ray(leaf, r, limit, &t, nrm)) return;
              if (have && !hit_less(t, (int64_t)leaf, best_t, best_leaf)) return;
              have = true; best_t = t; best_leaf = (int64_t)leaf; limit = t;
              for (int i = 0; i < D; ++i) normal_out[i] = nrm[i];
            });
  *leaf_out = best_leaf; *t_out = best_t;
  return have;
}

// Bvh::shapecast_nearest (bvh.rs:331-375): nodes tested by the probe-grown slab.
template <int D>
static bool bvh_cast_nearest(const Tree64<D>& tree, const Geom<D>& g, const Ray& r, const double* pmin, const double* pmax,
                             double max_toi, uint64_t mask, int64_t exclude, int64_t* leaf_out, double* t_out) {
  bool have = false;
  double limit = max_toi, best_t = 0.0;
  int64_t best_leaf = -1;
  tree.walk([&](const double* mn, const double* mx) { double te, tx; return g.broadphase(mn, mx, r, pmin, pmax, limit, &te, &tx); },
            [&](size_t leaf) {
              if (!g.accepts(leaf, mask, exclude)) return;
              double t;
              if (!g.test_cast(leaf, r, pmin, pmax, limit, &t)) return;
              if (have && !hit_less(t, (int64_t)leaf, best_t, best_leaf)) return;
              have = true; best_t = t; best_leaf = (int64_t)leaf; limit = t;
            });
  *leaf_out = best_leaf; *t_out = best_t;
  return have;
}

template <int D>
static void nearest_batch(int kindf, const double* prims, uint32_t n, const uint64_t* layers, uint64_t mask, int64_t exclude,
                          const double* rays, uint64_t nr, double max_toi, const double* tmax, int threads, int64_t* leaf,
                          double* t, double* point, double* normal, const Geom<D>* given = nullptr) {
  const Geom<D> g = given ? *given : geom_of<D>(kindf, n, prims, layers);
  const bool bvh = (kindf & FLAG_BVH) != 0;
  const Tree64<D> tree(bvh ? g : Geom<D>{g.kind, 0, prims, layers});
  parallel_for(nr, threads, [&](uint64_t i) {
    Ray r{rays + i * 2 * D, rays + i * 2 * D + D};
    double nrm[D];
    int64_t l; double tt;
    const double lim = tmax ? tmax[i] : max_toi;
    const bool have = bvh ? bvh_nearest<D>(tree, g, r, lim, mask, exclude, &l, &tt, nrm)
                          : naive_nearest<D>(g, r, lim, mask, exclude, &l, &tt, nrm);
    leaf[i] = have ? l : -1;
    t[i] = have ? tt : INFINITY;
    for (int k = 0; k < D; ++k) {
      if (point) point[i * D + k] = have ? r.o[k] + r.d[k] * tt : 0.0;   // ray.rs:33-35
      if (normal) normal[i * D + k] = have ? nrm[k] : 0.0;
    }
  });
}

template <int D>
static void cast_nearest_batch(int kindf, const double* prims, uint32_t n, const uint64_t* layers, uint64_t mask,
                               int64_t exclude, const double* casts, uint64_t nc, double max_toi, const double* tmax,
                               int threads, int64_t* leaf, double* t, double* point) {
  const Geom<D> g = geom_of<D>(kindf, n, prims, layers);
  const bool bvh = (kindf & FLAG_BVH) != 0;
  const Tree64<D> tree(bvh ? g : Geom<D>{g.kind, 0, prims, layers});
  parallel_for(nc, threads, [&](uint64_t i) {
    const double* c = casts + i * 4 * D;
    Ray r{c, c + D};
    int64_t l; double tt;
    const double lim = tmax ? tmax[i] : max_toi;
    const bool have = bvh ? bvh_cast_nearest<D>(tree, g, r, c + 2 * D, c + 3 * D, lim, mask, exclude, &l, &tt)
                          : naive_cast_nearest<D>(g, r, c + 2 * D, c + 3 * D, lim, mask, exclude, &l, &tt);
    leaf[i] = have ? l : -1;
    t[i] = have ? tt : INFINITY;
    for (int k = 0; k < D; ++k)
      if (point) point[i * D + k] = have ? r.o[k] + r.d[k] * tt : 0.0;   // ShapeCast::at
  });
}

// mode 0: rays, 1: overlap regions, 2: shape casts
template <int D>
static uint64_t all_batch(int kindf, const double* prims, uint32_t n, const uint64_t* layers, uint64_t mask, int64_t exclude,
                          const double* q, uint64_t nq, double max_toi, const double* tmax, int mode, int threads,
                          uint64_t* offsets, int64_t* leaf, double* t, uint64_t cap, const Geom<D>* given = nullptr) {
  const Geom<D> g = given ? *given : geom_of<D>(kindf, n, prims, layers);
  const bool bvh = (kindf & FLAG_BVH) != 0;
  const Tree64<D> tree(bvh ? g : Geom<D>{g.kind, 0, prims, layers});
  std::vector<std::vector<std::pair<double, int64_t>>> res(nq);
  auto by_hit = [](const std::pair<double, int64_t>& a, const std::pair<double, int64_t>& b) {
    return hit_less(a.first, a.second, b.first, b.second);
  };
  parallel_for(nq, threads, [&](uint64_t i) {
    if (bvh) {
      // bvh.rs:196-259 (all hits), :378-414 (cast all), :418-456 (overlap, by leaf)
      auto& out = res[i];
      if (mode == 2) {
        const double* c = q + i * 4 * D;
        Ray r{c, c + D};
        const double lim = tmax ? tmax[i] : max_toi;
        tree.walk([&](const double* mn, const double* mx) { double te, tx; return g.broadphase(mn, mx, r, c + 2 * D, c + 3 * D, lim, &te, &tx); },
                  [&](size_t lf) {
                    double tt;
                    if (g.accepts(lf, mask, exclude) && g.test_cast(lf, r, c + 2 * D, c + 3 * D, lim, &tt)) out.push_back({tt, (int64_t)lf});
                  });
        std::stable_sort(out.begin(), out.end(), by_hit);
      } else if (mode == 1) {
        const double *rmn = q + i * 2 * D, *rmx = rmn + D;
        tree.walk([&](const double* mn, const double* mx) {
                    for (int k = 0; k < D; ++k) if (mn[k] > rmx[k] || mx[k] < rmn[k]) return false;
                    return true;
                  },
                  [&](size_t lf) { if (g.accepts(lf, mask, exclude) && g.test_overlap(lf, rmn, rmx)) out.push_back({0.0, (int64_t)lf}); });
        std::stable_sort(out.begin(), out.end(), [](const std::pair<double, int64_t>& a, const std::pair<double, int64_t>& b) {
          return a.second < b.second;
        });
      } else {
        Ray r{q + i * 2 * D, q + i * 2 * D + D};
        tree.walk([&](const double* mn, const double* mx) { double te, tx; int ax; return ray_intersect<D>(mn, mx, r, max_toi, &te, &tx, &ax); },
                  [&](size_t lf) {
                    if (g.accepts(lf, mask, exclude)) g.list_hits(lf, r, max_toi, [&](double tt) { out.push_back({tt, (int64_t)lf}); });
                  });
        std::stable_sort(out.begin(), out.end(), by_hit);
      }
      return;
    }
    if (mode == 2) {
      const double* c = q + i * 4 * D;
      Ray r{c, c + D};
      naive_cast_all<D>(g, r, c + 2 * D, c + 3 * D, tmax ? tmax[i] : max_toi, mask, exclude, res[i]);
    } else if (mode == 1) {
      std::vector<int64_t> ids;
      naive_overlap<D>(g, q + i * 2 * D, q + i * 2 * D + D, mask, exclude, ids);
      for (int64_t id : ids) res[i].push_back({0.0, id});
    } else {
      Ray r{q + i * 2 * D, q + i * 2 * D + D};
      naive_all<D>(g, r, max_toi, mask, exclude, res[i]);
    }
  });
  uint64_t tot = 0;
  for (uint64_t i = 0; i < nq; ++i) { offsets[i] = tot; tot += res[i].size(); }
  offsets[nq] = tot;
  if (tot <= cap) {
    uint64_t k = 0;
    for (uint64_t i = 0; i < nq; ++i)
      for (auto& h : res[i]) { leaf[k] = h.second; if (t) t[k] = h.first; ++k; }
  }
  return tot;
}

// raycast_nearest_batch_parallel (bvh.rs:160-193): contiguous chunks, one per thread
// prims_build (may be NULL): the tree is built over these primitives and then
// refit to `prims` (Bvh::refit, bvh.rs:45-48), so it keeps their topology;
// *quality (may be NULL) receives its interior surface sum.
template <int D>
static void bvh_nearest_batch(int kindf, const double* prims, uint32_t n, const double* rays, uint64_t nr, double max_toi,
                              int threads, int64_t* leaf, double* t, uint64_t* counts, const double* prims_build = nullptr,
                              double* quality = nullptr) {
  const Geom<D> g = geom_of<D>(kindf, n, prims, nullptr);
  Tree64<D> tree(prims_build ? geom_of<D>(kindf, n, prims_build, nullptr) : g);
  if (prims_build) {
    std::vector<double> boxes((size_t)n * 2 * D);
    for (uint32_t i = 0; i < n; ++i) g.world_aabb(i, boxes.data() + (size_t)i * 2 * D, boxes.data() + (size_t)i * 2 * D + D);
    tree.refit(boxes);
  }
  if (quality) *quality = tree.quality();
  std::vector<uint64_t> nodes(std::max(threads, 1), 0), prs(std::max(threads, 1), 0);
  const int T = std::max(threads, 1);
  const uint64_t chunk = (nr + T - 1) / T;
  std::vector<std::thread> th;
  for (int w = 0; w < T; ++w) {
    th.emplace_back([&, w] {
      const uint64_t a = w * chunk, b = std::min<uint64_t>(nr, a + chunk);
      for (uint64_t i = a; i < b; ++i) {
        Ray r{rays + i * 2 * D, rays + i * 2 * D + D};
        int64_t l; double tt;
        const bool have = tree.nearest(g, r, max_toi, &l, &tt, &nodes[w], &prs[w]);
        leaf[i] = have ? l : -1;
        t[i] = have ? tt : INFINITY;
      }
    });
  }
  for (auto& x : th) x.join();
  counts[0] = counts[1] = 0;
  for (int w = 0; w < T; ++w) { counts[0] += nodes[w]; counts[1] += prs[w]; }
}

// Instances: R [D][D] row-major, t [D] each, and half [D] (boxes,
// two_level.rs:70-113: local box [-half, half]) or a radius (balls,
// shape_overlap.rs:288-345: local box [-r, r]^D, a ball at the local origin).
template <int D> struct Inst64 {
  uint64_t n;
  const double *rot, *trans, *half;   // half: [n][D] (boxes) or radii [n] (balls)
  bool guarded = false;   // the engine's SQ_HITS_GUARDED: a hit kept inside its world box's slab interval
  int kind = 0;           // 0 boxes, 1 balls
  double h(uint64_t i, int k) const { return kind ? half[i] : half[i * D + k]; }
  // Isometry::transform_aabb of the local box [-h, h]: centre R c + t, half extent |R| h
  void world_box(uint64_t i, double* mn, double* mx) const {
    const double *R = rot + i * D * D, *t = trans + i * D;
    double c[D], e[D];
    for (int k = 0; k < D; ++k) { const double hk = h(i, k), lo = hk * -1.0; c[k] = (lo + hk) * 0.5; e[k] = (hk - lo) * 0.5; }
    for (int a = 0; a < D; ++a) {
      double acc = 0.0;
      for (int b = 0; b < D; ++b) acc = acc + R[a * D + b] * c[b];
      const double rc = acc + t[a];
      double w = 0.0;
      for (int b = 0; b < D; ++b) w = w + std::fabs(R[a * D + b]) * e[b];
      mn[a] = rc - w; mx[a] = rc + w;
    }
  }
  // ray_ball (shape_overlap.rs:120-138) about centre c (NULL: the origin)
  static bool ray_ball(const double* o, const double* d, const double* c, double rad, double max_toi, double* t_out) {
    double oc[D];
    for (int k = 0; k < D; ++k) oc[k] = c ? o[k] - c[k] : o[k];
    const double b = dot<D>(oc, d);
    const double cc = dot<D>(oc, oc) - rad * rad;
    const double disc = b * b - cc;
    if (disc < 0.0) return false;
    const double sq = std::sqrt(disc);
    double t = -b - sq;
    if (t < 0.0) t = -b + sq;
    if (t < 0.0 || t > max_toi) return false;
    *t_out = t;
    return true;
  }
  // test_instance (tlas.rs:459-481): inverse transform (rotate_transpose of
  // o - t, and of d), then boxes: the local box slab with its 1e-9 parallel
  // threshold and t > 1e-6 entry rule; balls: ray_ball about the origin
  // (guarded: then the guard against the world box)
  bool test(uint64_t i, const Ray& r, double max_toi, double* t_out) const {
    const double *R = rot + i * D * D, *t = trans + i * D;
    double v[D], o[D], d[D];
    for (int j = 0; j < D; ++j) v[j] = r.o[j] - t[j];
    for (int a = 0; a < D; ++a) {
      double x = 0.0, y = 0.0;
      for (int j = 0; j < D; ++j) { x = x + R[j * D + a] * v[j]; y = y + R[j * D + a] * r.d[j]; }
      o[a] = x; d[a] = y;
    }
    double t_hit;
    if (kind) {
      if (!ray_ball(o, d, nullptr, half[i], max_toi, &t_hit)) return false;
    } else {
      double t_min = 0.0, t_max = max_toi;
      int axis = -1;
      for (int k = 0; k < D; ++k) {
        const double hk = half[i * D + k];
        if (std::fabs(d[k]) < 1.0e-9) {
          if (o[k] < -hk || o[k] > hk) return false;
        } else {
          const double inv = 1.0 / d[k];
          double t1 = (-hk - o[k]) * inv, t2 = (hk - o[k]) * inv;
          if (t1 > t2) std::swap(t1, t2);
          if (t1 > t_min) { t_min = t1; axis = k; }
          if (t2 < t_max) t_max = t2;
          if (t_min > t_max) return false;
        }
      }
      if (!(axis >= 0 && t_min > 1.0e-6 && t_min <= max_toi)) return false;
      t_hit = t_min;
    }
    if (guarded) {
      double mn[D], mx[D], te, tx; int ax;
      world_box(i, mn, mx);
      if (!ray_intersect<D>(mn, mx, r, max_toi, &te, &tx, &ax)) return false;
      if (!(te <= t_hit && t_hit <= tx)) return false;
    }
    *t_out = t_hit;
    return true;
  }
  // test_shape_cast: boxes keep the trait default (no contact, instance.rs:
  // 72-87); balls the cast ray against the ball of radius r + the probe's
  // max x about the centre (shape_overlap.rs:326-337, the built-in SPHERE
  // cast's probe radius), guarded: inside the probe-grown world box's slab
  // interval
  bool cast(uint64_t i, const Ray& r, const double* pmin, const double* pmax, double max_toi, double* t_out) const {
    if (!kind) return false;
    double t_hit;
    if (!ray_ball(r.o, r.d, trans + i * D, half[i] + pmax[0], max_toi, &t_hit)) return false;
    if (guarded) {
      double mn[D], mx[D], te, tx; int ax;
      world_box(i, mn, mx);
      for (int k = 0; k < D; ++k) { mn[k] = mn[k] - pmax[k]; mx[k] = mx[k] - pmin[k]; }
      if (!ray_intersect<D>(mn, mx, r, max_toi, &te, &tx, &ax)) return false;
      if (!(te <= t_hit && t_hit <= tx)) return false;
    }
    *t_out = t_hit;
    return true;
  }
  // test_overlap: boxes the trait default (true, instance.rs:89-98); balls
  // point_to_aabb(centre, region) <= r (shape_overlap.rs:59-72,339-341)
  bool overlap(uint64_t i, const double* rmn, const double* rmx) const {
    if (!kind) return true;
    const double* c = trans + i * D;
    double d2 = 0.0;
    for (int k = 0; k < D; ++k) {
      double v;
      if (c[k] < rmn[k]) v = rmn[k] - c[k];
      else if (c[k] > rmx[k]) v = c[k] - rmx[k];
      else v = 0.0;
      d2 = d2 + v * v;
    }
    return std::sqrt(d2) <= half[i];
  }
};

// flags: FLAG_GUARDED (the engine's guard), FLAG_BVH (the Tlas walk of
// tlas.rs:104-154 over the median-split tree of the world boxes, in the
// reference's order; else the naive instance scan of two_level.rs:126-167).
template <int D>
static void tlas_box_nearest(uint64_t n, const double* rot, const double* trans, const double* half, const double* rays,
                             uint64_t nr, double max_toi, int threads, int flags, int64_t* leaf, double* t, double* world) {
  Inst64<D> s{n, rot, trans, half};
  s.guarded = (flags & FLAG_GUARDED) != 0;
  std::vector<double> wb(n * 2 * D);
  for (uint64_t i = 0; i < n; ++i) s.world_box(i, wb.data() + i * 2 * D, wb.data() + i * 2 * D + D);
  if (world) std::copy(wb.begin(), wb.end(), world);
  const bool bvh = (flags & FLAG_BVH) != 0;
  const Tree64<D> tree(wb, bvh ? (uint32_t)n : 0u);
  parallel_for(nr, threads, [&](uint64_t q) {
    Ray r{rays + q * 2 * D, rays + q * 2 * D + D};
    int64_t bl = -1; double bt = 0.0;
    if (bvh) {
      double limit = max_toi;
      tree.walk([&](const double* mn, const double* mx) { double te, tx; int ax; return ray_intersect<D>(mn, mx, r, limit, &te, &tx, &ax); },
                [&](size_t i) {
                  double tt;
                  if (!s.test(i, r, limit, &tt)) return;
                  if (bl >= 0 && !hit_less(tt, (int64_t)i, bt, bl)) return;
                  bl = (int64_t)i; bt = tt; limit = tt;
                });
    } else {
      for (uint64_t i = 0; i < n; ++i) {   // two_level.rs:126-167
        double tt;
        if (!s.test(i, r, max_toi, &tt)) continue;
        if (bl < 0 || hit_less(tt, (int64_t)i, bt, bl)) { bl = (int64_t)i; bt = tt; }
      }
    }
    leaf[q] = bl;
    t[q] = bl >= 0 ? bt : INFINITY;
  });
}

// Every Tlas query over box (kind 0) or ball (kind 1) instances, flags as
// above.  mode 0: raycast_nearest (tlas.rs:104-154; naive: two_level.rs:
// 134-154); 1: raycast_all (tlas.rs:158-215; naive_all, two_level.rs:156-174)
// = raycast_crossings for these kinds (one crossing per instance, the
// test_ray_crossings_local default); 2: overlap (tlas.rs:393-432: every
// instance of a leaf node touching the region that passes test_overlap;
// naive: every instance whose world box touches the region and passes it,
// naive.rs:184-205's rule); 3: shapecast_nearest (tlas.rs:306-351); 4:
// shapecast_all (tlas.rs:353-391).  Rays [nq][2D], regions [nq][2D], casts
// [nq][4D] (origin, dir, probe min, max).  Modes 0 and 3 write leaf[nq],
// t[nq]; the others CSR (offsets[nq + 1]; leaf, t when the total fits cap)
// and return the total.
// build (may be NULL): {rot, trans, half} of the instances the tree was built
// over; it is then refit to the world boxes of the current ones (Tlas::refit,
// tlas.rs:49-54), keeping the build's topology; *quality (may be NULL)
// receives the tree's interior surface sum.
template <int D>
static uint64_t tlas_query(int kind, uint64_t n, const double* rot, const double* trans, const double* half, const double* q,
                           uint64_t nq, double max_toi, int mode, int threads, int flags, uint64_t* offsets, int64_t* leaf,
                           double* t, uint64_t cap, const double* const* build = nullptr, double* quality = nullptr) {
  Inst64<D> s{n, rot, trans, half};
  s.guarded = (flags & FLAG_GUARDED) != 0;
  s.kind = kind;
  std::vector<double> wb(n * 2 * D);
  for (uint64_t i = 0; i < n; ++i) s.world_box(i, wb.data() + i * 2 * D, wb.data() + i * 2 * D + D);
  const bool bvh = (flags & FLAG_BVH) != 0;
  std::vector<double> wb0;
  if (build) {
    Inst64<D> s0{n, build[0], build[1], build[2]};
    s0.kind = kind;
    wb0.resize(n * 2 * D);
    for (uint64_t i = 0; i < n; ++i) s0.world_box(i, wb0.data() + i * 2 * D, wb0.data() + i * 2 * D + D);
  }
  Tree64<D> tree(build ? wb0 : wb, bvh ? (uint32_t)n : 0u);
  if (build) tree.refit(wb);
  if (quality) *quality = tree.quality();
  std::vector<std::vector<std::pair<double, int64_t>>> res((mode == 0 || mode == 3) ? 0 : nq);
  auto by_hit = [](const std::pair<double, int64_t>& a, const std::pair<double, int64_t>& b) {
    return hit_less(a.first, a.second, b.first, b.second);
  };
  auto grown = [&](const double* mn, const double* mx, const Ray& r, const double* pmin, const double* pmax, double lim) {
    double gmn[D], gmx[D], te, tx; int ax;
    for (int k = 0; k < D; ++k) { gmn[k] = mn[k] - pmax[k]; gmx[k] = mx[k] - pmin[k]; }
    return ray_intersect<D>(gmn, gmx, r, lim, &te, &tx, &ax) && te <= lim;
  };
  parallel_for(nq, threads, [&](uint64_t i) {
    if (mode == 0 || mode == 3) {
      const int qs = mode == 0 ? 2 * D : 4 * D;
      const double* c = q + i * qs;
      Ray r{c, c + D};
      int64_t bl = -1; double bt = 0.0, limit = max_toi;
      auto take = [&](size_t k, double lim) {
        double tt;
        const bool h = mode == 0 ? s.test(k, r, lim, &tt) : s.cast(k, r, c + 2 * D, c + 3 * D, lim, &tt);
        if (!h) return;
        if (bl >= 0 && !hit_less(tt, (int64_t)k, bt, bl)) return;
        bl = (int64_t)k; bt = tt; limit = tt;
      };
      if (bvh) {
        tree.walk([&](const double* mn, const double* mx) {
                    if (mode == 3) return grown(mn, mx, r, c + 2 * D, c + 3 * D, limit);
                    double te, tx; int ax;
                    return ray_intersect<D>(mn, mx, r, limit, &te, &tx, &ax) && te <= limit;
                  },
                  [&](size_t k) { take(k, limit); });
      } else {
        for (uint64_t k = 0; k < n; ++k) take(k, max_toi);
      }
      leaf[i] = bl;
      t[i] = bl >= 0 ? bt : INFINITY;
      return;
    }
    auto& out = res[i];
    if (mode == 2) {
      const double *rmn = q + i * 2 * D, *rmx = rmn + D;
      auto touch = [&](const double* mn, const double* mx) {
        for (int k = 0; k < D; ++k) if (mn[k] > rmx[k] || mx[k] < rmn[k]) return false;
        return true;
      };
      if (bvh) tree.walk(touch, [&](size_t k) { if (s.overlap(k, rmn, rmx)) out.push_back({0.0, (int64_t)k}); });
      else
        for (uint64_t k = 0; k < n; ++k)
          if (touch(wb.data() + k * 2 * D, wb.data() + k * 2 * D + D) && s.overlap(k, rmn, rmx)) out.push_back({0.0, (int64_t)k});
      std::stable_sort(out.begin(), out.end(), [](const std::pair<double, int64_t>& a, const std::pair<double, int64_t>& b) {
        return a.second < b.second;
      });
    } else {
      const int qs = mode == 1 ? 2 * D : 4 * D;
      const double* c = q + i * qs;
      Ray r{c, c + D};
      auto take = [&](size_t k) {
        double tt;
        if (mode == 1 ? s.test(k, r, max_toi, &tt) : s.cast(k, r, c + 2 * D, c + 3 * D, max_toi, &tt))
          out.push_back({tt, (int64_t)k});
      };
      if (bvh)
        tree.walk([&](const double* mn, const double* mx) {
                    if (mode == 4) return grown(mn, mx, r, c + 2 * D, c + 3 * D, max_toi);
                    double te, tx; int ax;
                    return ray_intersect<D>(mn, mx, r, max_toi, &te, &tx, &ax);
                  },
                  take);
      else
        for (uint64_t k = 0; k < n; ++k) take(k);
      std::stable_sort(out.begin(), out.end(), by_hit);
    }
  });
  if (mode == 0 || mode == 3) return 0;
  uint64_t tot = 0;
  for (uint64_t i = 0; i < nq; ++i) { offsets[i] = tot; tot += res[i].size(); }
  offsets[nq] = tot;
  if (tot <= cap) {
    uint64_t k = 0;
    for (uint64_t i = 0; i < nq; ++i)
      for (auto& hh : res[i]) { leaf[k] = hh.second; if (t) t[k] = hh.first; ++k; }
  }
  return tot;
}

}  // namespace orc64

using namespace orc64;

#define ORC64_DISPATCH(D, CALL)                 \
  switch (D) {                                 \
    case 2: { constexpr int DD = 2; CALL; break; } \
    case 3: { constexpr int DD = 3; CALL; break; } \
    case 4: { constexpr int DD = 4; CALL; break; } \
    case 5: { constexpr int DD = 5; CALL; break; } \
    case 6: { constexpr int DD = 6; CALL; break; } \
    default: break;                            \
  }

// A consumer provider's queries (records [n][2D + pd]: world box, then the
// consumer's doubles): mode 0 nearest (FLAG_BVH: the reference's LIFO walk,
// else the naive scan), mode 1 all hits.  FLAG_GUARDED: hits kept inside
// their box's slab interval.
template <int D>
static uint64_t consumer_query(int flags, const double* records, uint32_t n, int pd, uint64_t fn, const double* rays,
                               uint64_t nr, double max_toi, int mode, int threads, int64_t* leaf, double* t,
                               uint64_t* offsets, uint64_t cap, uint64_t fnx = 0) {
  Geom<D> g{2, n, records, nullptr};
  g.guarded = (flags & FLAG_GUARDED) != 0;
  g.consumer = reinterpret_cast<ConsumerNarrow64>(fn);
  g.consumer_x = reinterpret_cast<ConsumerCrossings64>(fnx);
  g.consumer_pd = pd;
  const int kindf = 2 | (flags & FLAG_BVH);
  if (mode == 0) {
    nearest_batch<D>(kindf, records, n, nullptr, ~0ull, -1, rays, nr, max_toi, nullptr, threads, leaf, t, nullptr, nullptr, &g);
    return 0;
  }
  return all_batch<D>(kindf, records, n, nullptr, ~0ull, -1, rays, nr, max_toi, nullptr, 0, threads, offsets, leaf, t, cap, &g);
}

extern "C" {

void orc64_raycast_nearest(int D, int kind, const double* prims, uint32_t n, const uint64_t* layers, uint64_t mask,
                           int64_t exclude, const double* rays, uint64_t nr, double max_toi, const double* tmax, int threads,
                           int64_t* leaf, double* t, double* point, double* normal) {
  ORC64_DISPATCH(D, nearest_batch<DD>(kind, prims, n, layers, mask, exclude, rays, nr, max_toi, tmax, threads, leaf, t, point, normal));
}

uint64_t orc64_raycast_all(int D, int kind, const double* prims, uint32_t n, const uint64_t* layers, uint64_t mask,
                           int64_t exclude, const double* rays, uint64_t nr, double max_toi, int threads,
                           uint64_t* offsets, int64_t* leaf, double* t, uint64_t cap) {
  uint64_t tot = 0;
  ORC64_DISPATCH(D, tot = all_batch<DD>(kind, prims, n, layers, mask, exclude, rays, nr, max_toi, nullptr, 0, threads, offsets, leaf, t, cap));
  return tot;
}

uint64_t orc64_overlap(int D, int kind, const double* prims, uint32_t n, const uint64_t* layers, uint64_t mask,
                       int64_t exclude, const double* regions, uint64_t nq, int threads, uint64_t* offsets, int64_t* leaf,
                       uint64_t cap) {
  uint64_t tot = 0;
  ORC64_DISPATCH(D, tot = all_batch<DD>(kind, prims, n, layers, mask, exclude, regions, nq, 0.0, nullptr, 1, threads, offsets, leaf, nullptr, cap));
  return tot;
}

void orc64_shapecast_nearest(int D, int kind, const double* prims, uint32_t n, const uint64_t* layers, uint64_t mask,
                             int64_t exclude, const double* casts, uint64_t nc, double max_toi, const double* tmax,
                             int threads, int64_t* leaf, double* t, double* point) {
  ORC64_DISPATCH(D, cast_nearest_batch<DD>(kind, prims, n, layers, mask, exclude, casts, nc, max_toi, tmax, threads, leaf, t, point));
}

uint64_t orc64_shapecast_all(int D, int kind, const double* prims, uint32_t n, const uint64_t* layers, uint64_t mask,
                             int64_t exclude, const double* casts, uint64_t nc, double max_toi, const double* tmax,
                             int threads, uint64_t* offsets, int64_t* leaf, double* t, uint64_t cap) {
  uint64_t tot = 0;
  ORC64_DISPATCH(D, tot = all_batch<DD>(kind, prims, n, layers, mask, exclude, casts, nc, max_toi, tmax, 2, threads, offsets, leaf, t, cap));
  return tot;
}

// counts[2]: node boxes tested, primitives tested (summed over the rays)
void orc64_bvh_raycast_nearest(int D, int kind, const double* prims, uint32_t n, const double* rays, uint64_t nr,
                               double max_toi, int threads, int64_t* leaf, double* t, uint64_t* counts) {
  ORC64_DISPATCH(D, bvh_nearest_batch<DD>(kind, prims, n, rays, nr, max_toi, threads, leaf, t, counts));
}

// The same walk over the tree built on prims_build and refit to prims
// (Bvh::refit); quality: the refit tree's interior surface sum.
void orc64_bvh_refit_raycast_nearest(int D, int kind, const double* prims_build, const double* prims, uint32_t n,
                                     const double* rays, uint64_t nr, double max_toi, int threads, int64_t* leaf, double* t,
                                     uint64_t* counts, double* quality) {
  ORC64_DISPATCH(D, bvh_nearest_batch<DD>(kind, prims, n, rays, nr, max_toi, threads, leaf, t, counts, prims_build, quality));
}

void orc64_tlas_box_nearest(int D, uint64_t n, const double* rot, const double* trans, const double* half,
                            const double* rays, uint64_t nr, double max_toi, int threads, int flags, int64_t* leaf, double* t,
                            double* world) {
  ORC64_DISPATCH(D, tlas_box_nearest<DD>(n, rot, trans, half, rays, nr, max_toi, threads, flags, leaf, t, world));
}

// Tlas queries over box (kind 0) or ball (kind 1) instances: tlas_query.
uint64_t orc64_tlas_query(int D, int kind, uint64_t n, const double* rot, const double* trans, const double* half,
                          const double* q, uint64_t nq, double max_toi, int mode, int threads, int flags, uint64_t* offsets,
                          int64_t* leaf, double* t, uint64_t cap) {
  uint64_t tot = 0;
  ORC64_DISPATCH(D, tot = tlas_query<DD>(kind, n, rot, trans, half, q, nq, max_toi, mode, threads, flags, offsets, leaf, t, cap));
  return tot;
}

// tlas_query over the tree built on the instances {rot0, trans0, half0} and
// refit to {rot, trans, half} (Tlas::refit); quality: its interior surface sum.
uint64_t orc64_tlas_refit_query(int D, int kind, uint64_t n, const double* rot0, const double* trans0, const double* half0,
                                const double* rot, const double* trans, const double* half, const double* q, uint64_t nq,
                                double max_toi, int mode, int threads, int flags, uint64_t* offsets, int64_t* leaf, double* t,
                                uint64_t cap, double* quality) {
  uint64_t tot = 0;
  const double* build[3] = {rot0, trans0, half0};
  ORC64_DISPATCH(D, tot = tlas_query<DD>(kind, n, rot, trans, half, q, nq, max_toi, mode, threads, flags, offsets, leaf, t, cap,
                                         build, quality));
  return tot;
}

uint64_t orc64_consumer_query(int D, int flags, const double* records, uint32_t n, int pd, uint64_t fn, const double* rays,
                              uint64_t nr, double max_toi, int mode, int threads, int64_t* leaf, double* t, uint64_t* offsets,
                              uint64_t cap) {
  uint64_t tot = 0;
  ORC64_DISPATCH(D, tot = consumer_query<DD>(flags, records, n, pd, fn, rays, nr, max_toi, mode, threads, leaf, t, offsets, cap));
  return tot;
}

// Bvh::raycast_crossings over a consumer whose crossings function is fnx
// (test_ray_crossings): CSR lists as orc64_consumer_query's mode 1.
uint64_t orc64_consumer_crossings(int D, int flags, const double* records, uint32_t n, int pd, uint64_t fn, uint64_t fnx,
                                  const double* rays, uint64_t nr, double max_toi, int threads, int64_t* leaf, double* t,
                                  uint64_t* offsets, uint64_t cap) {
  uint64_t tot = 0;
  ORC64_DISPATCH(D, tot = consumer_query<DD>(flags, records, n, pd, fn, rays, nr, max_toi, 1, threads, leaf, t, offsets, cap, fnx));
  return tot;
}

void orc64_tree_stats(int D, int kind, const double* prims, uint32_t n, uint64_t* out5, double* quality) {
  ORC64_DISPATCH(D, tree_stats<DD>(prims, kind & 0xFF, n, out5, quality));
}

}  // extern "C"
