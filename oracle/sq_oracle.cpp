// sq_oracle.cpp -- CPU oracle for the batched BVH query path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain C++, the CPU path of
// grimandgreedy/spatial-query (crate `spatial-query` v0.1.0) so that the HIP
// engine can be checked against it bit for bit.  Only `tests/`, the smoke check
// in `__graft_entry__.py` and the `cpu_baseline` leg of `bench.py` may load it;
// the product library never links, loads or calls it.
//
// Every function follows the reference operation for operation, in the same
// order, with FP contraction disabled (built with -ffp-contract=off), so each
// f32 result carries the same bits the Rust code would produce:
//   * f32::min/max                 -> fminf/fmaxf (minNum/maxNum)
//   * f32::total_cmp               -> total_key() below (sign-magnitude order)
//   * slice::sort_by               -> std::stable_sort
//   * `u64 as f32`                 -> (float)u64 (round to nearest even)
// Citations are `path:line` relative to the reference repository root with
// `sq/` = `crates/spatial-query/`.
//
// Built-in providers (the engine's narrow phases; the reference core has none):
//   BOX    (N1) toi = Aabb::ray_intersect(leaf box).t_min, sq/src/maths/aabb.rs:118-148
//   SPHERE (N2) sq/tests/differential.rs:61-82 (exit root when inside)
//   TRI    (N3) Moeller-Trumbore crates/viewport-lib-query/src/lib.rs:279-306,
//               plus the `t > max_toi` reject of sq/src/query/geometry.rs:37-39
// By default every provider is plain, as the reference's: the results are the
// reference's.  A geometry or instance set marked `guarded` restates the
// engine's SQ_HITS_GUARDED mode instead: SPHERE, TRI and consumer hits (and
// instance hits, against the instance's world box) are kept only when their
// toi lies inside the slab interval of their own AABB, which makes every hit
// lie inside every ancestor box, so the nearest result no longer depends on
// traversal order (BVH == naive then holds exactly).  The *_PLAIN kinds are
// the plain providers under names of their own.
//
// Parity pins for each part live in tests/test_oracle.py (closed forms from the
// reference tests, BVH == naive, refit == rebuild, parallel == sequential,
// brute force on tiny inputs).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

namespace orc {

typedef float Scalar;
static const float F32_EPSILON = 1.1920929e-7f;  // f32::EPSILON
static const uint32_t LEAF_SIZE = 4;             // sq/src/accel/tree.rs:12

// ---------------------------------------------------------------------------
// L0 maths: sq/src/maths/point.rs, aabb.rs, ray.rs, isometry.rs
// ---------------------------------------------------------------------------

template <int D> struct Point {
  float v[D];
  float operator[](int i) const { return v[i]; }
  float& operator[](int i) { return v[i]; }
  static Point splat(float s) { Point p; for (int i = 0; i < D; ++i) p.v[i] = s; return p; }
  static Point zero() { return splat(0.0f); }
};

// point.rs:40-56 (f32::min / f32::max)
template <int D> Point<D> pmin(Point<D> a, Point<D> b) { for (int i = 0; i < D; ++i) a.v[i] = fminf(a.v[i], b.v[i]); return a; }
template <int D> Point<D> pmax(Point<D> a, Point<D> b) { for (int i = 0; i < D; ++i) a.v[i] = fmaxf(a.v[i], b.v[i]); return a; }
// point.rs:100-134
template <int D> Point<D> padd(Point<D> a, Point<D> b) { for (int i = 0; i < D; ++i) a.v[i] = a.v[i] + b.v[i]; return a; }
template <int D> Point<D> psub(Point<D> a, Point<D> b) { for (int i = 0; i < D; ++i) a.v[i] = a.v[i] - b.v[i]; return a; }
template <int D> Point<D> pmul(Point<D> a, float s) { for (int i = 0; i < D; ++i) a.v[i] = a.v[i] * s; return a; }
// point.rs:60-66: accumulate from 0.0, left to right, no FMA.
template <int D> float pdot(Point<D> a, Point<D> b) {
  float acc = 0.0f;
  for (int i = 0; i < D; ++i) acc = acc + a.v[i] * b.v[i];
  return acc;
}
// point.rs:82-89
template <int D> Point<D> normalize_or_zero(Point<D> a) {
  float len = sqrtf(pdot(a, a));
  if (len > F32_EPSILON) return pmul(a, 1.0f / len);
  return Point<D>::zero();
}
template <int D> bool peq(const Point<D>& a, const Point<D>& b) {
  for (int i = 0; i < D; ++i) if (!(a.v[i] == b.v[i])) return false;
  return true;
}

template <int D> struct Ray {
  Point<D> origin, dir;
  // ray.rs:33-35
  Point<D> at(float t) const { return padd(origin, pmul(dir, t)); }
};
// ray.rs:18-23
template <int D> Ray<D> ray_new(Point<D> o, Point<D> d) { Ray<D> r; r.origin = o; r.dir = normalize_or_zero(d); return r; }

template <int D> struct Aabb {
  Point<D> min, max;
  // aabb.rs:18-21
  static Aabb empty() { Aabb b; b.min = Point<D>::splat(INFINITY); b.max = Point<D>::splat(-INFINITY); return b; }
  // aabb.rs:38-43
  Aabb unite(const Aabb& o) const { Aabb b; b.min = pmin(min, o.min); b.max = pmax(max, o.max); return b; }
  // aabb.rs:47-52
  Aabb union_point(const Point<D>& p) const { Aabb b; b.min = pmin(min, p); b.max = pmax(max, p); return b; }
  // aabb.rs:56-58
  Point<D> center() const { return pmul(padd(min, max), 0.5f); }
  // aabb.rs:62-64
  Point<D> extents() const { return psub(max, min); }
  // aabb.rs:68-77: first maximum wins ties.
  int longest_axis() const {
    Point<D> e = extents();
    int best = 0;
    for (int i = 1; i < D; ++i) if (e.v[i] > e.v[best]) best = i;
    return best;
  }
  // aabb.rs:84-97
  float half_surface() const {
    Point<D> e = extents();
    float sum = 0.0f;
    for (int i = 0; i < D; ++i) {
      float prod = 1.0f;
      for (int j = 0; j < D; ++j) if (j != i) prod = prod * e.v[j];
      sum = sum + prod;
    }
    return sum;
  }
  // aabb.rs:101-108: touching counts.
  bool intersects(const Aabb& o) const {
    for (int i = 0; i < D; ++i)
      if (min.v[i] > o.max.v[i] || max.v[i] < o.min.v[i]) return false;
    return true;
  }
  // aabb.rs:118-148: slab test.  Returns false on a miss.
  bool ray_intersect(const Ray<D>& ray, float max_toi, float* out_tmin, float* out_tmax, int* entry_axis = nullptr) const {
    float t_min = 0.0f;
    float t_max = max_toi;
    int axis = -1;
    for (int i = 0; i < D; ++i) {
      float o = ray.origin.v[i];
      float d = ray.dir.v[i];
      if (fabsf(d) < F32_EPSILON) {
        if (o < min.v[i] || o > max.v[i]) return false;
      } else {
        float inv = 1.0f / d;
        float t1 = (min.v[i] - o) * inv;
        float t2 = (max.v[i] - o) * inv;
        if (t1 > t2) { float tmp = t1; t1 = t2; t2 = tmp; }
        if (t1 > t_min) { t_min = t1; axis = i; }
        if (t2 < t_max) t_max = t2;
        if (t_min > t_max) return false;
      }
    }
    if (out_tmin) *out_tmin = t_min;
    if (out_tmax) *out_tmax = t_max;
    if (entry_axis) *entry_axis = axis;
    return true;
  }
};

// isometry.rs:20-124 (row-major rotation, rotate / rotate_transpose accumulate
// from 0.0 left to right).
template <int D> struct Isometry {
  float rot[D][D];
  Point<D> trans;
  Point<D> rotate(Point<D> v) const {
    Point<D> out;
    for (int i = 0; i < D; ++i) {
      float acc = 0.0f;
      for (int j = 0; j < D; ++j) acc = acc + rot[i][j] * v.v[j];
      out.v[i] = acc;
    }
    return out;
  }
  Point<D> rotate_transpose(Point<D> v) const {
    Point<D> out;
    for (int i = 0; i < D; ++i) {
      float acc = 0.0f;
      for (int j = 0; j < D; ++j) acc = acc + rot[j][i] * v.v[j];
      out.v[i] = acc;
    }
    return out;
  }
  Point<D> transform_point(Point<D> p) const { return padd(rotate(p), trans); }          // :58-60
  Point<D> transform_vector(Point<D> v) const { return rotate(v); }                      // :64-66
  Point<D> inverse_transform_point(Point<D> p) const { return rotate_transpose(psub(p, trans)); }  // :70-72
  Point<D> inverse_transform_vector(Point<D> v) const { return rotate_transpose(v); }    // :76-78
  // :84-97
  Aabb<D> transform_aabb(const Aabb<D>& local) const {
    Point<D> c = transform_point(local.center());
    Point<D> h = pmul(local.extents(), 0.5f);
    Point<D> wh;
    for (int i = 0; i < D; ++i) {
      float acc = 0.0f;
      for (int j = 0; j < D; ++j) acc = acc + fabsf(rot[i][j]) * h.v[j];
      wh.v[i] = acc;
    }
    Aabb<D> b; b.min = psub(c, wh); b.max = padd(c, wh);
    return b;
  }
};

// f32::total_cmp as a signed integer key (sign-magnitude order).
static inline int32_t total_key(float f) {
  int32_t i; std::memcpy(&i, &f, 4);
  i ^= (int32_t)(((uint32_t)(i >> 31)) >> 1);
  return i;
}
static inline int total_cmp(float a, float b) {
  int32_t x = total_key(a), y = total_key(b);
  return x < y ? -1 : (x > y ? 1 : 0);
}

// ---------------------------------------------------------------------------
// L1 query types: sq/src/query/{hit,filter,shapecast}.rs
// ---------------------------------------------------------------------------

struct Filter {           // filter.rs:12-29
  uint64_t layer_mask;
  int64_t exclude;        // -1 = None
};

template <int D> struct LeafHit {   // hit.rs:20-30
  float toi;
  Point<D> normal;
  int64_t sub;            // -1 = None
};

template <int D> struct Hit {       // hit.rs:55-71
  int64_t id;
  int64_t leaf;
  float toi;
  Point<D> point;
  Point<D> normal;
  int64_t sub;
};

// hit.rs:77-84: toi by total_cmp, then leaf.
template <int D> int hit_ordering(const Hit<D>& a, const Hit<D>& b) {
  int c = total_cmp(a.toi, b.toi);
  if (c != 0) return c;
  return a.leaf < b.leaf ? -1 : (a.leaf > b.leaf ? 1 : 0);
}

template <int D> struct ShapeCast {  // shapecast.rs:26-73
  Point<D> origin, dir;
  Aabb<D> aabb;
  Point<D> at(float t) const { return padd(origin, pmul(dir, t)); }
  // shapecast.rs:70-73: slab of the ray against target grown by the probe box.
  bool broadphase(const Aabb<D>& target, float max_toi, float* te, float* tx) const {
    Aabb<D> grown; grown.min = psub(target.min, aabb.max); grown.max = psub(target.max, aabb.min);
    Ray<D> r; r.origin = origin; r.dir = dir;
    return grown.ray_intersect(r, max_toi, te, tx);
  }
};

// ---------------------------------------------------------------------------
// Built-in providers (QueryGeometry, sq/src/query/geometry.rs:19-98)
// ---------------------------------------------------------------------------

enum Kind {
  KIND_BOX = 0,              // N1
  KIND_SPHERE = 1,           // N2 (box-guarded when the geometry is `guarded`)
  KIND_TRI = 2,              // N3, D = 3 (box-guarded when the geometry is `guarded`)
  KIND_SPHERE_PLAIN = 10,    // differential.rs:61-82 exactly (exit root when inside)
  KIND_SPHERE_PLAIN_NOEXIT = 11,  // conformance.rs:277-291 / lib.rs doc example
  KIND_TRI_PLAIN = 12,       // viewport-lib-query MT + max_toi reject, unguarded
  KIND_SHELL = 13,           // crossings.rs:37-111: both sphere roots are crossings
  KIND_CONSUMER = 20,        // a consumer's test_ray (geometry.rs:33-43) through a C function (box-guarded when `guarded`)
};

// A consumer narrow phase: true with *toi on a hit of the primitive's floats.
typedef int (*ConsumerNarrow)(const float* prim, const float* origin, const float* dir, float max_toi, float* toi);
// A consumer's crossings (test_ray_crossings, geometry.rs:45-63): calls
// emit(ctx, toi) once per crossing of the ray with the primitive.
typedef void (*ConsumerCrossings)(const float* prim, const float* origin, const float* dir, float max_toi,
                                  void (*emit)(void*, float), void* ctx);
// A consumer's swept-box contact (test_shape_cast, geometry.rs:65-77) and exact
// overlap test (test_overlap, geometry.rs:79-87).
typedef int (*ConsumerCast)(const float* prim, const float* origin, const float* dir, const float* probe_min,
                            const float* probe_max, float max_toi, float* toi);
typedef int (*ConsumerOverlap)(const float* prim, const float* region_min, const float* region_max);

template <int D> struct Geometry {
  int kind;
  uint32_t n;
  std::vector<float> prims;      // BOX: min D, max D; SPHERE: center D, radius; TRI: v0 v1 v2
  std::vector<uint64_t> layers;  // empty = every leaf on every layer
  ConsumerNarrow consumer = nullptr;   // KIND_CONSUMER: records are [world box 2D][consumer_pf floats]
  ConsumerCrossings consumer_x = nullptr;   // its crossings, when it reports more than the nearest hit
  ConsumerCast consumer_cast = nullptr;     // its shape-cast contact (none without)
  ConsumerOverlap consumer_ov = nullptr;    // its exact overlap test (the box test alone without)
  int consumer_pf = 0;
  bool guarded = false;                     // SPHERE / TRI / CONSUMER: the engine's SQ_HITS_GUARDED restated
  int stride() const {
    if (kind == KIND_CONSUMER) return 2 * D + consumer_pf;
    return kind == KIND_BOX ? 2 * D : (kind == KIND_TRI || kind == KIND_TRI_PLAIN) ? 9 : D + 1;
  }
  bool is_ball() const { return kind == KIND_SPHERE || kind == KIND_SPHERE_PLAIN || kind == KIND_SPHERE_PLAIN_NOEXIT || kind == KIND_SHELL; }
  const float* p(size_t leaf) const { return prims.data() + leaf * stride(); }

  size_t leaf_count() const { return n; }
  int64_t id(size_t leaf) const { return (int64_t)leaf; }

  Aabb<D> world_aabb(size_t leaf) const {
    const float* q = p(leaf);
    Aabb<D> b;
    if (kind == KIND_BOX || kind == KIND_CONSUMER) {
      for (int i = 0; i < D; ++i) { b.min.v[i] = q[i]; b.max.v[i] = q[D + i]; }
    } else if (kind == KIND_TRI || kind == KIND_TRI_PLAIN) {
      // The AABB of the three vertices, folded with union_point from EMPTY.
      b = Aabb<D>::empty();
      for (int k = 0; k < 3; ++k) { Point<D> v; for (int i = 0; i < D; ++i) v.v[i] = q[3 * k + i]; b = b.union_point(v); }
    } else {
      // Ball: Aabb::new(c - splat(r), c + splat(r)), differential.rs:55-59.
      Point<D> c; for (int i = 0; i < D; ++i) c.v[i] = q[i];
      float r = q[D];
      b.min = psub(c, Point<D>::splat(r)); b.max = padd(c, Point<D>::splat(r));
    }
    return b;
  }

  // filter semantic of the built-in kinds: a leaf is accepted when it shares a
  // layer bit with the mask and is not the excluded id.
  bool accepts(size_t leaf, const Filter& f) const {
    if (f.exclude >= 0 && (int64_t)leaf == f.exclude) return false;
    uint64_t lay = layers.empty() ? ~0ull : layers[leaf];
    return (lay & f.layer_mask) != 0;
  }

  // The guard (only for a `guarded` geometry): toi must lie in the slab
  // interval of the leaf's own AABB.
  bool guard(size_t leaf, const Ray<D>& ray, float max_toi, float t) const {
    if (!guarded) return true;
    float te, tx;
    if (!world_aabb(leaf).ray_intersect(ray, max_toi, &te, &tx)) return false;
    return te <= t && t <= tx;
  }

  bool ball(const Ray<D>& ray, Point<D> c, float r, float max_toi, bool take_exit, float* out_t) const {
    Point<D> oc = psub(ray.origin, c);
    float b = pdot(oc, ray.dir);
    float cc = pdot(oc, oc) - r * r;
    float disc = b * b - cc;
    if (disc < 0.0f) return false;
    float sq = sqrtf(disc);
    float t = -b - sq;
    if (take_exit && t < 0.0f) t = -b + sq;
    if (t < 0.0f || t > max_toi) return false;
    *out_t = t;
    return true;
  }

  // crossings.rs:49-61: the two parameters where the unbounded ray meets the sphere.
  bool roots(const Ray<D>& ray, Point<D> c, float r, float* t0, float* t1) const {
    Point<D> oc = psub(ray.origin, c);
    float b = pdot(oc, ray.dir);
    float disc = b * b - (pdot(oc, oc) - r * r);
    if (disc < 0.0f) return false;
    float sq = sqrtf(disc);
    *t0 = -b - sq; *t1 = -b + sq;
    return true;
  }

  bool test_ray(size_t leaf, const Ray<D>& ray, float max_toi, LeafHit<D>* out) const {
    const float* q = p(leaf);
    if (kind == KIND_CONSUMER) {
      float t;
      if (!consumer(q + 2 * D, ray.origin.v, ray.dir.v, max_toi, &t)) return false;
      if (!(t >= 0.0f && t <= max_toi)) return false;   // the test_ray contract, geometry.rs:37-39
      if (!guard(leaf, ray, max_toi, t)) return false;
      out->toi = t;
      out->normal = Point<D>::zero();
      out->sub = -1;
      return true;
    }
    if (kind == KIND_SHELL) {   // crossings.rs:76-86
      Point<D> c; for (int i = 0; i < D; ++i) c.v[i] = q[i];
      float t0, t1;
      if (!roots(ray, c, q[D], &t0, &t1)) return false;
      float t = t0 >= 0.0f ? t0 : t1;
      if (t < 0.0f || t > max_toi) return false;
      out->toi = t; out->normal = normalize_or_zero(psub(ray.at(t), c)); out->sub = -1;
      return true;
    }
    if (kind == KIND_BOX) {
      float te, tx; int axis;
      if (!world_aabb(leaf).ray_intersect(ray, max_toi, &te, &tx, &axis)) return false;
      out->toi = te;
      out->normal = Point<D>::zero();
      if (axis >= 0) out->normal.v[axis] = ray.dir.v[axis] > 0.0f ? -1.0f : 1.0f;
      out->sub = -1;
      return true;
    }
    if (kind == KIND_SPHERE || kind == KIND_SPHERE_PLAIN || kind == KIND_SPHERE_PLAIN_NOEXIT) {
      Point<D> c; for (int i = 0; i < D; ++i) c.v[i] = q[i];
      float t;
      if (!ball(ray, c, q[D], max_toi, kind != KIND_SPHERE_PLAIN_NOEXIT, &t)) return false;
      if (kind == KIND_SPHERE && !guard(leaf, ray, max_toi, t)) return false;   // (a no-op unless guarded)
      out->toi = t;
      out->normal = normalize_or_zero(psub(ray.at(t), c));
      out->sub = (int64_t)leaf;
      return true;
    }
    // Moeller-Trumbore, viewport-lib-query/src/lib.rs:281-306.  Cross products
    // are written out; dots go through Point::dot (same bits for nonzero sums).
    if (D != 3) return false;
    Point<D> v0, v1, v2;
    for (int i = 0; i < 3; ++i) { v0.v[i] = q[i]; v1.v[i] = q[3 + i]; v2.v[i] = q[6 + i]; }
    Point<D> e1 = psub(v1, v0), e2 = psub(v2, v0);
    const Point<D>& d = ray.dir;
    Point<D> pv = cross(d, e2);
    float det = pdot(e1, pv);
    if (fabsf(det) < 1e-8f) return false;
    float inv = 1.0f / det;
    Point<D> tvec = psub(ray.origin, v0);
    float u = pdot(tvec, pv) * inv;
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    Point<D> qv = cross(tvec, e1);
    float v = pdot(d, qv) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = pdot(e2, qv) * inv;
    if (!(t > 1e-6f)) return false;
    if (t > max_toi) return false;
    if (kind == KIND_TRI && !guard(leaf, ray, max_toi, t)) return false;
    out->toi = t;
    Point<D> nrm = normalize_or_zero(cross(e1, e2));
    if (pdot(nrm, d) > 0.0f) nrm = pmul(nrm, -1.0f);
    out->normal = nrm;
    out->sub = (int64_t)leaf;
    return true;
  }

  static Point<D> cross(const Point<D>& a, const Point<D>& b) {
    Point<D> r = Point<D>::zero();
    r.v[0] = a.v[1] * b.v[2] - a.v[2] * b.v[1];
    r.v[1] = a.v[2] * b.v[0] - a.v[0] * b.v[2];
    r.v[2] = a.v[0] * b.v[1] - a.v[1] * b.v[0];
    return r;
  }

  // Crossings (geometry.rs:45-55 default = the single test_ray hit).  Boxes and
  // balls are convex: report entry and exit.
  template <class F> void test_ray_crossings(size_t leaf, const Ray<D>& ray, float max_toi, F&& out) const {
    if (kind == KIND_CONSUMER && consumer_x) {
      // each crossing kept under the test_ray contract and the box guard, as
      // the consumer's test_ray hits; emitted in the consumer's order
      std::vector<float> ts;
      consumer_x(p(leaf) + 2 * D, ray.origin.v, ray.dir.v, max_toi,
                 [](void* c, float t) { static_cast<std::vector<float>*>(c)->push_back(t); }, &ts);
      for (float t : ts) {
        if (!(t >= 0.0f && t <= max_toi) || !guard(leaf, ray, max_toi, t)) continue;
        LeafHit<D> h; h.toi = t; h.normal = Point<D>::zero(); h.sub = -1;
        out(h);
      }
      return;
    }
    if (kind == KIND_SHELL) {   // crossings.rs:87-110
      const float* q = p(leaf);
      Point<D> c; for (int i = 0; i < D; ++i) c.v[i] = q[i];
      float t0, t1;
      if (!roots(ray, c, q[D], &t0, &t1)) return;
      float ts[2] = {t0, t1};
      for (float t : ts) {
        if (t >= 0.0f && t <= max_toi) {
          LeafHit<D> h; h.toi = t; h.normal = normalize_or_zero(psub(ray.at(t), c)); h.sub = -1;
          out(h);
        }
      }
      return;
    }
    if (kind == KIND_BOX) {
      // entry and exit of the slab interval, each kept when inside [0, max_toi]
      float te, tx; int axis;
      if (!world_aabb(leaf).ray_intersect(ray, max_toi, &te, &tx, &axis)) return;
      LeafHit<D> h; h.sub = -1;
      h.toi = te; h.normal = Point<D>::zero();
      if (axis >= 0) h.normal.v[axis] = ray.dir.v[axis] > 0.0f ? -1.0f : 1.0f;
      out(h);
      if (tx > te) {
        // The exit parameter is reported only when it is the box's real exit
        // (not the max_toi clamp).
        float te2, tx2;
        world_aabb(leaf).ray_intersect(ray, INFINITY, &te2, &tx2);
        if (tx2 <= max_toi) { h.toi = tx2; h.normal = Point<D>::zero(); out(h); }
      }
      return;
    }
    LeafHit<D> h;
    if (test_ray(leaf, ray, max_toi, &h)) out(h);
  }

  // Shape cast narrow (N4 for boxes; summed-radius ball for spheres,
  // shape_overlap.rs:104-118).  Triangles keep the default None.
  bool test_shape_cast(size_t leaf, const ShapeCast<D>& cast, float max_toi, LeafHit<D>* out) const {
    if (kind == KIND_CONSUMER) {
      // the contact under the test_shape_cast contract (and, guarded, inside
      // the probe-grown slab interval of the primitive's world box)
      if (!consumer_cast) return false;
      float t;
      if (!consumer_cast(p(leaf) + 2 * D, cast.origin.v, cast.dir.v, cast.aabb.min.v, cast.aabb.max.v, max_toi, &t)) return false;
      if (!(t >= 0.0f && t <= max_toi)) return false;
      float te, tx;
      if (guarded && !cast.broadphase(world_aabb(leaf), max_toi, &te, &tx)) return false;
      if (guarded && !(te <= t && t <= tx)) return false;
      out->toi = t; out->normal = Point<D>::zero(); out->sub = -1;
      return true;
    }
    if (kind == KIND_BOX) {
      float te, tx;
      if (!cast.broadphase(world_aabb(leaf), max_toi, &te, &tx)) return false;
      out->toi = te; out->normal = Point<D>::zero(); out->sub = -1;
      return true;
    }
    if (is_ball()) {
      const float* q = p(leaf);
      Point<D> c; for (int i = 0; i < D; ++i) c.v[i] = q[i];
      float sum = q[D] + cast.aabb.max.v[0];
      Ray<D> r; r.origin = cast.origin; r.dir = cast.dir;
      float t;
      if (!ball(r, c, sum, max_toi, true, &t)) return false;
      if (kind == KIND_SPHERE && guarded) {
        float te, tx;
        if (!cast.broadphase(world_aabb(leaf), max_toi, &te, &tx)) return false;
        if (!(te <= t && t <= tx)) return false;
      }
      out->toi = t;
      out->normal = normalize_or_zero(psub(cast.at(t), c));
      out->sub = (int64_t)leaf;
      return true;
    }
    return false;
  }

  // Exact overlap: box intersects; ball = nearest point of the region within r
  // (shape_overlap.rs:60-72,121-124); triangle = its AABB.
  bool test_overlap(size_t leaf, const Aabb<D>& region) const {
    if (kind == KIND_BOX || kind == KIND_TRI || kind == KIND_TRI_PLAIN) return world_aabb(leaf).intersects(region);
    if (kind == KIND_CONSUMER) {
      if (!world_aabb(leaf).intersects(region)) return false;
      return !consumer_ov || consumer_ov(p(leaf) + 2 * D, region.min.v, region.max.v);
    }
    const float* q = p(leaf);
    float d2 = 0.0f;
    for (int i = 0; i < D; ++i) {
      float pi = q[i], v;
      if (pi < region.min.v[i]) v = region.min.v[i] - pi;
      else if (pi > region.max.v[i]) v = pi - region.max.v[i];
      else v = 0.0f;
      d2 = d2 + v * v;
    }
    return sqrtf(d2) <= q[D];
  }
};

// ---------------------------------------------------------------------------
// L2 tree: sq/src/accel/tree.rs
// ---------------------------------------------------------------------------

template <int D> struct Node {   // tree.rs:20-26
  Aabb<D> bounds;
  uint32_t child_a, child_b, prim_count;
  bool is_leaf() const { return prim_count > 0; }
};

template <int D>
static uint32_t build_recursive(std::vector<Node<D>>& nodes, std::vector<uint32_t>& indices,
                                const std::vector<Aabb<D>>& bounds, const std::vector<Point<D>>& centroids,
                                uint32_t start, uint32_t end) {
  // tree.rs:120-123
  Aabb<D> node_bounds = Aabb<D>::empty();
  for (uint32_t k = start; k < end; ++k) node_bounds = node_bounds.unite(bounds[indices[k]]);
  uint32_t node_index = (uint32_t)nodes.size();
  uint32_t count = end - start;
  // tree.rs:129-134: reserve the slot in pre-order.
  Node<D> nd; nd.bounds = node_bounds; nd.child_a = start; nd.child_b = 0; nd.prim_count = count;
  nodes.push_back(nd);
  if (count <= LEAF_SIZE) return node_index;  // :136-138
  // :140-144
  Aabb<D> cb = Aabb<D>::empty();
  for (uint32_t k = start; k < end; ++k) cb = cb.union_point(centroids[indices[k]]);
  int axis = cb.longest_axis();
  if (cb.extents().v[axis] <= F32_EPSILON) return node_index;  // :145-148
  // :152-153: stable sort by total_cmp on the centroid axis.
  std::stable_sort(indices.begin() + start, indices.begin() + end, [&](uint32_t a, uint32_t b) {
    return total_key(centroids[a].v[axis]) < total_key(centroids[b].v[axis]);
  });
  uint32_t mid = start + count / 2;  // :154
  uint32_t ca = build_recursive(nodes, indices, bounds, centroids, start, mid);
  uint32_t cbi = build_recursive(nodes, indices, bounds, centroids, mid, end);
  nodes[node_index].child_a = ca;
  nodes[node_index].child_b = cbi;
  nodes[node_index].prim_count = 0;
  return node_index;
}

// tree.rs:39-49
template <int D>
static void tree_build(const std::vector<Aabb<D>>& bounds, std::vector<Node<D>>& nodes, std::vector<uint32_t>& indices) {
  uint32_t n = (uint32_t)bounds.size();
  indices.resize(n);
  for (uint32_t i = 0; i < n; ++i) indices[i] = i;
  nodes.clear();
  if (n == 0) return;
  std::vector<Point<D>> centroids(n);
  for (uint32_t i = 0; i < n; ++i) centroids[i] = bounds[i].center();
  nodes.reserve(2 * (size_t)n);
  build_recursive(nodes, indices, bounds, centroids, 0, n);
}

// tree.rs:54-75
template <int D>
static void tree_refit(std::vector<Node<D>>& nodes, const std::vector<uint32_t>& indices, const std::vector<Aabb<D>>& prim_bounds) {
  for (size_t ni = nodes.size(); ni-- > 0;) {
    Node<D> node = nodes[ni];
    if (node.is_leaf()) {
      Aabb<D> b = Aabb<D>::empty();
      for (uint32_t k = node.child_a; k < node.child_a + node.prim_count; ++k) b = b.unite(prim_bounds[indices[k]]);
      nodes[ni].bounds = b;
    } else {
      nodes[ni].bounds = nodes[node.child_a].bounds.unite(nodes[node.child_b].bounds);
    }
  }
}

// tree.rs:81-99 -> (leaf_count, prim_count, max_depth)
template <int D>
static void tree_structural(const std::vector<Node<D>>& nodes, uint64_t* leaves, uint64_t* prims, uint64_t* depth_out) {
  *leaves = *prims = *depth_out = 0;
  if (nodes.empty()) return;
  std::vector<uint64_t> depth(nodes.size(), 0);
  for (size_t ni = nodes.size(); ni-- > 0;) {
    const Node<D>& nd = nodes[ni];
    if (nd.is_leaf()) { (*leaves)++; *prims += nd.prim_count; depth[ni] = 1; }
    else depth[ni] = 1 + std::max(depth[nd.child_a], depth[nd.child_b]);
  }
  *depth_out = depth[0];
}

// tree.rs:104-110 (Iterator::sum over f32 accumulates from 0.0 in node order)
template <int D>
static float tree_interior_surface_sum(const std::vector<Node<D>>& nodes) {
  float s = 0.0f;
  for (const Node<D>& nd : nodes) if (!nd.is_leaf()) s = s + nd.bounds.half_surface();
  return s;
}

struct Stats { uint64_t nodes_visited, aabb_tests, narrow_tests, hits; };  // stats.rs:39-50

// ---------------------------------------------------------------------------
// L2 Bvh: sq/src/accel/bvh.rs
// ---------------------------------------------------------------------------

template <int D> struct Bvh {
  std::vector<Node<D>> nodes;
  std::vector<uint32_t> prim_indices;

  // bvh.rs:31-37
  void build(const Geometry<D>& g) {
    std::vector<Aabb<D>> b(g.leaf_count());
    for (size_t i = 0; i < b.size(); ++i) b[i] = g.world_aabb(i);
    tree_build(b, nodes, prim_indices);
  }
  // bvh.rs:45-48
  void refit(const Geometry<D>& g) {
    std::vector<Aabb<D>> b(g.leaf_count());
    for (size_t i = 0; i < b.size(); ++i) b[i] = g.world_aabb(i);
    tree_refit(nodes, prim_indices, b);
  }

  Hit<D> assemble(const Geometry<D>& g, const Ray<D>& ray, size_t leaf, const LeafHit<D>& lh) const {  // naive.rs:17-31
    Hit<D> h; h.id = g.id(leaf); h.leaf = (int64_t)leaf; h.toi = lh.toi; h.point = ray.at(lh.toi); h.normal = lh.normal; h.sub = lh.sub;
    return h;
  }

  // bvh.rs:73-124
  bool raycast_nearest(const Geometry<D>& g, const Ray<D>& ray, float max_toi, const Filter& f, Hit<D>* best_out, Stats* st) const {
    if (nodes.empty()) return false;
    bool have = false;
    Hit<D> best;
    float limit = max_toi;
    std::vector<uint32_t> stack(1, 0u);
    while (!stack.empty()) {
      uint32_t ni = stack.back(); stack.pop_back();
      const Node<D> node = nodes[ni];
      if (st) { st->nodes_visited++; st->aabb_tests++; }
      float te, tx;
      if (!node.bounds.ray_intersect(ray, limit, &te, &tx) || !(te <= limit)) continue;
      if (node.is_leaf()) {
        for (uint32_t k = node.child_a; k < node.child_a + node.prim_count; ++k) {
          size_t leaf = prim_indices[k];
          if (!g.accepts(leaf, f)) continue;
          if (st) st->narrow_tests++;
          LeafHit<D> lh;
          if (g.test_ray(leaf, ray, limit, &lh)) {
            if (st) st->hits++;
            Hit<D> hit = assemble(g, ray, leaf, lh);
            if (have && hit_ordering(best, hit) <= 0) {
            } else {
              limit = hit.toi; best = hit; have = true;
            }
          }
        }
      } else {
        stack.push_back(node.child_a);
        stack.push_back(node.child_b);
      }
    }
    if (have) *best_out = best;
    return have;
  }

  // bvh.rs:218-259 (all) and :264-328 (crossings)
  void raycast_all(const Geometry<D>& g, const Ray<D>& ray, float max_toi, const Filter& f, bool crossings, std::vector<Hit<D>>& hits, Stats* st) const {
    hits.clear();
    if (nodes.empty()) return;
    std::vector<uint32_t> stack(1, 0u);
    while (!stack.empty()) {
      uint32_t ni = stack.back(); stack.pop_back();
      const Node<D> node = nodes[ni];
      if (st) { st->nodes_visited++; st->aabb_tests++; }
      float te, tx;
      if (!node.bounds.ray_intersect(ray, max_toi, &te, &tx)) continue;
      if (node.is_leaf()) {
        for (uint32_t k = node.child_a; k < node.child_a + node.prim_count; ++k) {
          size_t leaf = prim_indices[k];
          if (!g.accepts(leaf, f)) continue;
          if (st) st->narrow_tests++;
          if (crossings) {
            g.test_ray_crossings(leaf, ray, max_toi, [&](const LeafHit<D>& lh) { if (st) st->hits++; hits.push_back(assemble(g, ray, leaf, lh)); });
          } else {
            LeafHit<D> lh;
            if (g.test_ray(leaf, ray, max_toi, &lh)) { if (st) st->hits++; hits.push_back(assemble(g, ray, leaf, lh)); }
          }
        }
      } else {
        stack.push_back(node.child_a);
        stack.push_back(node.child_b);
      }
    }
    std::stable_sort(hits.begin(), hits.end(), [](const Hit<D>& a, const Hit<D>& b) { return hit_ordering(a, b) < 0; });
  }

  Hit<D> assemble_shape(const Geometry<D>& g, const ShapeCast<D>& c, size_t leaf, const LeafHit<D>& lh) const {  // naive.rs:114-128
    Hit<D> h; h.id = g.id(leaf); h.leaf = (int64_t)leaf; h.toi = lh.toi; h.point = c.at(lh.toi); h.normal = lh.normal; h.sub = lh.sub;
    return h;
  }

  // bvh.rs:331-375
  bool shapecast_nearest(const Geometry<D>& g, const ShapeCast<D>& cast, float max_toi, const Filter& f, Hit<D>* best_out) const {
    if (nodes.empty()) return false;
    bool have = false; Hit<D> best; float limit = max_toi;
    std::vector<uint32_t> stack(1, 0u);
    while (!stack.empty()) {
      uint32_t ni = stack.back(); stack.pop_back();
      const Node<D> node = nodes[ni];
      float te, tx;
      if (!cast.broadphase(node.bounds, limit, &te, &tx) || !(te <= limit)) continue;
      if (node.is_leaf()) {
        for (uint32_t k = node.child_a; k < node.child_a + node.prim_count; ++k) {
          size_t leaf = prim_indices[k];
          if (!g.accepts(leaf, f)) continue;
          LeafHit<D> lh;
          if (g.test_shape_cast(leaf, cast, limit, &lh)) {
            Hit<D> hit = assemble_shape(g, cast, leaf, lh);
            if (have && hit_ordering(best, hit) <= 0) {
            } else { limit = hit.toi; best = hit; have = true; }
          }
        }
      } else { stack.push_back(node.child_a); stack.push_back(node.child_b); }
    }
    if (have) *best_out = best;
    return have;
  }

  // bvh.rs:378-414
  void shapecast_all(const Geometry<D>& g, const ShapeCast<D>& cast, float max_toi, const Filter& f, std::vector<Hit<D>>& hits) const {
    hits.clear();
    if (nodes.empty()) return;
    std::vector<uint32_t> stack(1, 0u);
    while (!stack.empty()) {
      uint32_t ni = stack.back(); stack.pop_back();
      const Node<D> node = nodes[ni];
      float te, tx;
      if (!cast.broadphase(node.bounds, max_toi, &te, &tx)) continue;
      if (node.is_leaf()) {
        for (uint32_t k = node.child_a; k < node.child_a + node.prim_count; ++k) {
          size_t leaf = prim_indices[k];
          if (!g.accepts(leaf, f)) continue;
          LeafHit<D> lh;
          if (g.test_shape_cast(leaf, cast, max_toi, &lh)) hits.push_back(assemble_shape(g, cast, leaf, lh));
        }
      } else { stack.push_back(node.child_a); stack.push_back(node.child_b); }
    }
    std::stable_sort(hits.begin(), hits.end(), [](const Hit<D>& a, const Hit<D>& b) { return hit_ordering(a, b) < 0; });
  }

  // bvh.rs:418-456
  void overlap(const Geometry<D>& g, const Aabb<D>& region, const Filter& f, std::vector<int64_t>& out) const {
    out.clear();
    if (nodes.empty()) return;
    std::vector<uint32_t> stack(1, 0u);
    while (!stack.empty()) {
      uint32_t ni = stack.back(); stack.pop_back();
      const Node<D> node = nodes[ni];
      if (!node.bounds.intersects(region)) continue;
      if (node.is_leaf()) {
        for (uint32_t k = node.child_a; k < node.child_a + node.prim_count; ++k) {
          size_t leaf = prim_indices[k];
          if (!g.accepts(leaf, f)) continue;
          if (g.test_overlap(leaf, region)) out.push_back((int64_t)leaf);
        }
      } else { stack.push_back(node.child_a); stack.push_back(node.child_b); }
    }
    std::stable_sort(out.begin(), out.end());  // sort_by_key(leaf)
  }

  // Order-independent minimal work of one ray at limit t: nodes tested and
  // primitives tested when every node on the path passes the slab at t.
  void required_work(const Ray<D>& ray, float t, uint64_t* n_nodes, uint64_t* n_prims) const {
    *n_nodes = 0; *n_prims = 0;
    if (nodes.empty()) return;
    *n_nodes = 1;
    std::vector<uint32_t> stack(1, 0u);
    while (!stack.empty()) {
      uint32_t ni = stack.back(); stack.pop_back();
      const Node<D>& node = nodes[ni];
      float te, tx;
      if (!node.bounds.ray_intersect(ray, t, &te, &tx)) continue;
      if (node.is_leaf()) { *n_prims += node.prim_count; }
      else { *n_nodes += 2; stack.push_back(node.child_a); stack.push_back(node.child_b); }
    }
  }
};

// ---------------------------------------------------------------------------
// naive scans: sq/src/query/naive.rs
// ---------------------------------------------------------------------------

template <int D>
static bool naive_nearest(const Geometry<D>& g, const Ray<D>& ray, float max_toi, const Filter& f, Hit<D>* out) {  // :34-62
  bool have = false; Hit<D> best; float limit = max_toi;
  for (size_t leaf = 0; leaf < g.leaf_count(); ++leaf) {
    if (!g.accepts(leaf, f)) continue;
    float te, tx;
    if (!g.world_aabb(leaf).ray_intersect(ray, limit, &te, &tx)) continue;
    LeafHit<D> lh;
    if (g.test_ray(leaf, ray, limit, &lh)) {
      Hit<D> h; h.id = g.id(leaf); h.leaf = (int64_t)leaf; h.toi = lh.toi; h.point = ray.at(lh.toi); h.normal = lh.normal; h.sub = lh.sub;
      if (have && hit_ordering(best, h) <= 0) {} else { limit = h.toi; best = h; have = true; }
    }
  }
  if (have) *out = best;
  return have;
}

template <int D>
static void naive_all(const Geometry<D>& g, const Ray<D>& ray, float max_toi, const Filter& f, bool crossings, std::vector<Hit<D>>& hits) {  // :65-109
  hits.clear();
  for (size_t leaf = 0; leaf < g.leaf_count(); ++leaf) {
    if (!g.accepts(leaf, f)) continue;
    float te, tx;
    if (!g.world_aabb(leaf).ray_intersect(ray, max_toi, &te, &tx)) continue;
    auto push = [&](const LeafHit<D>& lh) {
      Hit<D> h; h.id = g.id(leaf); h.leaf = (int64_t)leaf; h.toi = lh.toi; h.point = ray.at(lh.toi); h.normal = lh.normal; h.sub = lh.sub;
      hits.push_back(h);
    };
    if (crossings) g.test_ray_crossings(leaf, ray, max_toi, push);
    else { LeafHit<D> lh; if (g.test_ray(leaf, ray, max_toi, &lh)) push(lh); }
  }
  std::stable_sort(hits.begin(), hits.end(), [](const Hit<D>& a, const Hit<D>& b) { return hit_ordering(a, b) < 0; });
}

template <int D>
static bool naive_shapecast_nearest(const Geometry<D>& g, const ShapeCast<D>& c, float max_toi, const Filter& f, Hit<D>* out) {  // :131-158
  bool have = false; Hit<D> best; float limit = max_toi;
  for (size_t leaf = 0; leaf < g.leaf_count(); ++leaf) {
    if (!g.accepts(leaf, f)) continue;
    float te, tx;
    if (!c.broadphase(g.world_aabb(leaf), limit, &te, &tx)) continue;
    LeafHit<D> lh;
    if (g.test_shape_cast(leaf, c, limit, &lh)) {
      Hit<D> h; h.id = g.id(leaf); h.leaf = (int64_t)leaf; h.toi = lh.toi; h.point = c.at(lh.toi); h.normal = lh.normal; h.sub = lh.sub;
      if (have && hit_ordering(best, h) <= 0) {} else { limit = h.toi; best = h; have = true; }
    }
  }
  if (have) *out = best;
  return have;
}

template <int D>
static void naive_shapecast_all(const Geometry<D>& g, const ShapeCast<D>& c, float max_toi, const Filter& f, std::vector<Hit<D>>& hits) {  // :161-181
  hits.clear();
  for (size_t leaf = 0; leaf < g.leaf_count(); ++leaf) {
    if (!g.accepts(leaf, f)) continue;
    float te, tx;
    if (!c.broadphase(g.world_aabb(leaf), max_toi, &te, &tx)) continue;
    LeafHit<D> lh;
    if (g.test_shape_cast(leaf, c, max_toi, &lh)) {
      Hit<D> h; h.id = g.id(leaf); h.leaf = (int64_t)leaf; h.toi = lh.toi; h.point = c.at(lh.toi); h.normal = lh.normal; h.sub = lh.sub;
      hits.push_back(h);
    }
  }
  std::stable_sort(hits.begin(), hits.end(), [](const Hit<D>& a, const Hit<D>& b) { return hit_ordering(a, b) < 0; });
}

template <int D>
static void naive_overlap(const Geometry<D>& g, const Aabb<D>& region, const Filter& f, std::vector<int64_t>& out) {  // :184-205
  out.clear();
  for (size_t leaf = 0; leaf < g.leaf_count(); ++leaf) {
    if (!g.accepts(leaf, f)) continue;
    if (!g.world_aabb(leaf).intersects(region)) continue;
    if (g.test_overlap(leaf, region)) out.push_back((int64_t)leaf);
  }
}

// ---------------------------------------------------------------------------
// L2 Tlas: sq/src/accel/tlas.rs over built-in instance kinds
// ---------------------------------------------------------------------------

enum InstKind {
  INST_BOX = 0,          // local box [-half, half], two_level.rs:70-113 (ray_local_box)
  INST_MESH = 1,         // uniformly scaled triangle mesh with its own BVH (D = 3)
  INST_BOX_GUARDED = 2,  // INST_BOX in a `guarded` set (orc_inst_new's shorthand)
  INST_BALL = 3,         // ball of radius r at the local origin (shape_overlap.rs:288-345 BallInstances)
};

template <int D> struct Instances {
  int kind;
  std::vector<Isometry<D>> xf;
  std::vector<Point<D>> halves;        // INST_BOX
  std::vector<uint32_t> mesh_of;       // INST_MESH
  std::vector<float> scale;            // INST_MESH: uniform scale of the mesh
  std::vector<float> radius;           // INST_BALL
  std::vector<std::unique_ptr<Geometry<D>>> meshes;
  std::vector<std::unique_ptr<Bvh<D>>> mesh_bvh;
  std::vector<Aabb<D>> mesh_box;       // AABB of each mesh's vertices
  // SQ_HITS_GUARDED restated: every instance hit is kept only inside its world
  // box's slab interval, and a mesh's triangles are the guarded kind
  bool guarded = false;

  size_t instance_count() const { return xf.size(); }
  int64_t id(size_t i) const { return (int64_t)i; }

  // The AABB of instance i in its own (scaled) local space.
  bool is_box() const { return kind == INST_BOX || kind == INST_BOX_GUARDED; }
  Aabb<D> local_aabb(size_t i) const {
    if (is_box()) { Aabb<D> b; b.min = pmul(halves[i], -1.0f); b.max = halves[i]; return b; }
    if (kind == INST_BALL) {                                          // shape_overlap.rs:320-322
      Aabb<D> b;
      for (int k = 0; k < D; ++k) { b.min.v[k] = -radius[i]; b.max.v[k] = radius[i]; }
      return b;
    }
    const Aabb<D>& m = mesh_box[mesh_of[i]];
    Aabb<D> b; b.min = pmul(m.min, scale[i]); b.max = pmul(m.max, scale[i]);
    return b;
  }

  // two_level.rs:81-113 for boxes.  Meshes: the local ray is moved into the
  // unscaled mesh frame by o/s, d/s (t is unchanged) and answered by the
  // mesh's own Bvh::raycast_nearest (bvh.rs:51-124) over the triangle kind
  // (guarded in a `guarded` set).
  bool test_ray_local(size_t i, const Ray<D>& ray, float max_toi, LeafHit<D>* out) const {
    if (is_box()) {
      const Point<D>& half = halves[i];
      float t_min = 0.0f, t_max = max_toi; int axis = -1;
      for (int k = 0; k < D; ++k) {
        float o = ray.origin.v[k], d = ray.dir.v[k];
        if (fabsf(d) < 1.0e-9f) {
          if (o < -half.v[k] || o > half.v[k]) return false;
        } else {
          float inv = 1.0f / d;
          float t1 = (-half.v[k] - o) * inv;
          float t2 = (half.v[k] - o) * inv;
          if (t1 > t2) { float tmp = t1; t1 = t2; t2 = tmp; }
          if (t1 > t_min) { t_min = t1; axis = k; }
          if (t2 < t_max) t_max = t2;
          if (t_min > t_max) return false;
        }
      }
      if (!(axis >= 0 && t_min > 1.0e-6f && t_min <= max_toi)) return false;
      out->toi = t_min;
      out->normal = Point<D>::zero();
      out->normal.v[axis] = ray.dir.v[axis] > 0.0f ? -1.0f : 1.0f;
      out->sub = -1;
      return true;
    }
    if (kind == INST_BALL) {                                          // shape_overlap.rs:120-138 (ray_ball)
      float t;
      if (!ray_ball(ray, Point<D>::zero(), radius[i], max_toi, &t)) return false;
      out->toi = t; out->normal = normalize_or_zero(ray.at(t)); out->sub = -1;
      return true;
    }
    float inv_s = 1.0f / scale[i];
    Ray<D> mr; mr.origin = pmul(ray.origin, inv_s); mr.dir = pmul(ray.dir, inv_s);
    Hit<D> h;
    Filter all{~0ull, -1};
    if (!mesh_bvh[mesh_of[i]]->raycast_nearest(*meshes[mesh_of[i]], mr, max_toi, all, &h, nullptr)) return false;
    out->toi = h.toi; out->normal = h.normal; out->sub = h.leaf;
    return true;
  }

  // ray_ball (shape_overlap.rs:120-138): the entry root, else the exit root.
  static bool ray_ball(const Ray<D>& ray, const Point<D>& c, float r, float max_toi, float* t_out) {
    Point<D> oc = psub(ray.origin, c);
    float b = pdot(oc, ray.dir);
    float cc = pdot(oc, oc) - r * r;
    float disc = b * b - cc;
    if (disc < 0.0f) return false;
    float sq = sqrtf(disc);
    float t = -b - sq;
    if (t < 0.0f) t = -b + sq;
    if (t < 0.0f || t > max_toi) return false;
    *t_out = t;
    return true;
  }

  // InstancedGeometry::test_shape_cast (instance.rs:72-87): the swept ball
  // probe (radius = the cast box's max x) against ball instances
  // (shape_overlap.rs:326-337: the ray against the radius sum); None for the
  // other kinds, the trait's default.
  bool test_shape_cast(size_t i, const ShapeCast<D>& cast, float max_toi, LeafHit<D>* out) const {
    if (kind != INST_BALL) return false;
    const Point<D> c = xf[i].trans;
    Ray<D> r; r.origin = cast.origin; r.dir = cast.dir;
    float t;
    if (!ray_ball(r, c, radius[i] + cast.aabb.max.v[0], max_toi, &t)) return false;
    out->toi = t; out->normal = normalize_or_zero(psub(cast.at(t), c)); out->sub = -1;
    return true;
  }

  // InstancedGeometry::test_overlap (instance.rs:89-98): ball instances test
  // the distance from the centre to the region (shape_overlap.rs:59-72,
  // 339-341); the other kinds keep the default (true).
  bool test_overlap(size_t i, const Aabb<D>& region) const {
    if (kind != INST_BALL) return true;
    const Point<D> p = xf[i].trans;
    float d2 = 0.0f;
    for (int k = 0; k < D; ++k) {
      float v = p.v[k] < region.min.v[k] ? region.min.v[k] - p.v[k] : (p.v[k] > region.max.v[k] ? p.v[k] - region.max.v[k] : 0.0f);
      d2 += v * v;
    }
    return sqrtf(d2) <= radius[i];
  }

  bool accepts(size_t, const Filter& f) const { return f.layer_mask != 0; }
};

template <int D> struct Tlas {
  std::vector<Node<D>> nodes;
  std::vector<uint32_t> idx;
  std::vector<Aabb<D>> world;
  float baseline = 0.0f;

  void build(const Instances<D>& s) {                 // tlas.rs:35-45
    world.resize(s.instance_count());
    for (size_t i = 0; i < world.size(); ++i) world[i] = s.xf[i].transform_aabb(s.local_aabb(i));
    tree_build(world, nodes, idx);
    baseline = tree_interior_surface_sum(nodes);
  }
  void refit(const Instances<D>& s) {                 // tlas.rs:49-54
    for (size_t i = 0; i < world.size(); ++i) world[i] = s.xf[i].transform_aabb(s.local_aabb(i));
    tree_refit(nodes, idx, world);
  }
  bool needs_rebuild(float factor) const {            // tlas.rs:58-60
    return baseline > 0.0f && tree_interior_surface_sum(nodes) > factor * baseline;
  }

  // tlas.rs:459-481.  A `guarded` set adds the box guard against the
  // instance's world AABB, which makes the two-level result traversal-order
  // independent (the engine's SQ_HITS_GUARDED).
  bool test_instance(const Instances<D>& s, size_t i, const Ray<D>& ray, float max_toi, Hit<D>* out) const {
    const Isometry<D>& iso = s.xf[i];
    Ray<D> local; local.origin = iso.inverse_transform_point(ray.origin); local.dir = iso.inverse_transform_vector(ray.dir);
    LeafHit<D> lh;
    if (!s.test_ray_local(i, local, max_toi, &lh)) return false;
    if (s.guarded) {
      float te, tx;
      if (!world[i].ray_intersect(ray, max_toi, &te, &tx)) return false;
      if (!(te <= lh.toi && lh.toi <= tx)) return false;
    }
    out->id = s.id(i); out->leaf = (int64_t)i; out->toi = lh.toi; out->point = ray.at(lh.toi);
    out->normal = iso.transform_vector(lh.normal); out->sub = lh.sub;
    return true;
  }

  // tlas.rs:104-154
  bool raycast_nearest(const Instances<D>& s, const Ray<D>& ray, float max_toi, const Filter& f, Hit<D>* best_out) const {
    if (nodes.empty()) return false;
    bool have = false; Hit<D> best; float limit = max_toi;
    std::vector<uint32_t> stack(1, 0u);
    while (!stack.empty()) {
      uint32_t ni = stack.back(); stack.pop_back();
      const Node<D> node = nodes[ni];
      float te, tx;
      if (!node.bounds.ray_intersect(ray, limit, &te, &tx) || !(te <= limit)) continue;
      if (node.is_leaf()) {
        for (uint32_t k = node.child_a; k < node.child_a + node.prim_count; ++k) {
          size_t i = idx[k];
          if (!s.accepts(i, f)) continue;
          Hit<D> hit;
          if (test_instance(s, i, ray, limit, &hit)) {
            if (have && hit_ordering(best, hit) <= 0) {} else { limit = hit.toi; best = hit; have = true; }
          }
        }
      } else { stack.push_back(node.child_a); stack.push_back(node.child_b); }
    }
    if (have) *best_out = best;
    return have;
  }

  // shapecast_instance (tlas.rs:437-452); a `guarded` set keeps the contact
  // only inside the swept probe's slab interval against the instance's world
  // AABB (broadphase), which makes the result order independent.
  bool shapecast_instance(const Instances<D>& s, size_t i, const ShapeCast<D>& cast, float max_toi, Hit<D>* out) const {
    LeafHit<D> lh;
    if (!s.test_shape_cast(i, cast, max_toi, &lh)) return false;
    float te, tx;
    if (s.guarded && !cast.broadphase(world[i], max_toi, &te, &tx)) return false;
    if (s.guarded && !(te <= lh.toi && lh.toi <= tx)) return false;
    out->id = s.id(i); out->leaf = (int64_t)i; out->toi = lh.toi; out->point = cast.at(lh.toi);
    out->normal = lh.normal; out->sub = lh.sub;
    return true;
  }

  // tlas.rs:306-351
  bool shapecast_nearest(const Instances<D>& s, const ShapeCast<D>& cast, float max_toi, const Filter& f, Hit<D>* best_out) const {
    if (nodes.empty()) return false;
    bool have = false; Hit<D> best; float limit = max_toi;
    std::vector<uint32_t> stack(1, 0u);
    while (!stack.empty()) {
      uint32_t ni = stack.back(); stack.pop_back();
      const Node<D> node = nodes[ni];
      float te, tx;
      if (!cast.broadphase(node.bounds, limit, &te, &tx) || !(te <= limit)) continue;
      if (node.is_leaf()) {
        for (uint32_t k = node.child_a; k < node.child_a + node.prim_count; ++k) {
          size_t i = idx[k];
          if (!s.accepts(i, f)) continue;
          Hit<D> hit;
          if (shapecast_instance(s, i, cast, limit, &hit)) {
            if (have && hit_ordering(best, hit) <= 0) {} else { limit = hit.toi; best = hit; have = true; }
          }
        }
      } else { stack.push_back(node.child_a); stack.push_back(node.child_b); }
    }
    if (have) *best_out = best;
    return have;
  }

  // tlas.rs:353-391
  void shapecast_all(const Instances<D>& s, const ShapeCast<D>& cast, float max_toi, const Filter& f, std::vector<Hit<D>>& hits) const {
    hits.clear();
    if (nodes.empty()) return;
    std::vector<uint32_t> stack(1, 0u);
    while (!stack.empty()) {
      uint32_t ni = stack.back(); stack.pop_back();
      const Node<D> node = nodes[ni];
      float te, tx;
      if (!cast.broadphase(node.bounds, max_toi, &te, &tx)) continue;
      if (node.is_leaf()) {
        for (uint32_t k = node.child_a; k < node.child_a + node.prim_count; ++k) {
          size_t i = idx[k];
          if (!s.accepts(i, f)) continue;
          Hit<D> hit;
          if (shapecast_instance(s, i, cast, max_toi, &hit)) hits.push_back(hit);
        }
      } else { stack.push_back(node.child_a); stack.push_back(node.child_b); }
    }
    std::stable_sort(hits.begin(), hits.end(), [](const Hit<D>& a, const Hit<D>& b) { return hit_ordering(a, b) < 0; });
  }

  // tlas.rs:158-215 (raycast_all)
  void raycast_all(const Instances<D>& s, const Ray<D>& ray, float max_toi, const Filter& f, std::vector<Hit<D>>& hits) const {
    hits.clear();
    if (nodes.empty()) return;
    std::vector<uint32_t> stack(1, 0u);
    while (!stack.empty()) {
      uint32_t ni = stack.back(); stack.pop_back();
      const Node<D> node = nodes[ni];
      float te, tx;
      if (!node.bounds.ray_intersect(ray, max_toi, &te, &tx)) continue;
      if (node.is_leaf()) {
        for (uint32_t k = node.child_a; k < node.child_a + node.prim_count; ++k) {
          size_t i = idx[k];
          if (!s.accepts(i, f)) continue;
          Hit<D> hit;
          if (test_instance(s, i, ray, max_toi, &hit)) hits.push_back(hit);
        }
      } else { stack.push_back(node.child_a); stack.push_back(node.child_b); }
    }
    std::stable_sort(hits.begin(), hits.end(), [](const Hit<D>& a, const Hit<D>& b) { return hit_ordering(a, b) < 0; });
  }

  // tlas.rs:393-432: the instance's test_overlap once its world AABB meets the
  // region (the default, true, for boxes and meshes).
  void overlap(const Instances<D>& s, const Aabb<D>& region, const Filter& f, std::vector<int64_t>& out) const {
    out.clear();
    if (nodes.empty()) return;
    std::vector<uint32_t> stack(1, 0u);
    while (!stack.empty()) {
      uint32_t ni = stack.back(); stack.pop_back();
      const Node<D> node = nodes[ni];
      if (!node.bounds.intersects(region)) continue;
      if (node.is_leaf()) {
        for (uint32_t k = node.child_a; k < node.child_a + node.prim_count; ++k) {
          size_t i = idx[k];
          if (!s.accepts(i, f)) continue;
          if (!s.test_overlap(i, region)) continue;
          out.push_back((int64_t)i);
        }
      } else { stack.push_back(node.child_a); stack.push_back(node.child_b); }
    }
    std::stable_sort(out.begin(), out.end());
  }
};

// two_level.rs:126-167: naive scan over all instances with the same tie-break.
template <int D>
static bool naive_tlas_nearest(const Tlas<D>& t, const Instances<D>& s, const Ray<D>& ray, float max_toi, Hit<D>* out) {
  bool have = false; Hit<D> best;
  for (size_t i = 0; i < s.instance_count(); ++i) {
    Hit<D> h;
    if (t.test_instance(s, i, ray, max_toi, &h)) {
      if (!have || hit_ordering(h, best) < 0) { best = h; have = true; }
    }
  }
  if (have) *out = best;
  return have;
}

// ---------------------------------------------------------------------------
// Generators: SplitMix64, sq/tests/conformance.rs:9-25
// ---------------------------------------------------------------------------

struct Rng {
  uint64_t s;
  uint64_t next_u64() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  float range(float lo, float hi) {
    float u = (float)(next_u64() >> 11) / (float)(1ull << 53);
    return lo + u * (hi - lo);
  }
};

}  // namespace orc

// ===========================================================================
// C API for the test harness (ctypes).  D is a runtime argument dispatched to
// the template instantiations 1..8.
// ===========================================================================

using namespace orc;

struct OGeom { int D; void* p; };
struct OBvh { int D; void* p; };
struct OInst { int D; void* p; void* tlas; };

#define ORC_DISPATCH(D, ...)                                   \
  switch (D) {                                                 \
    case 1: { const int DD = 1; __VA_ARGS__; } break;                 \
    case 2: { const int DD = 2; __VA_ARGS__; } break;                 \
    case 3: { const int DD = 3; __VA_ARGS__; } break;                 \
    case 4: { const int DD = 4; __VA_ARGS__; } break;                 \
    case 5: { const int DD = 5; __VA_ARGS__; } break;                 \
    case 6: { const int DD = 6; __VA_ARGS__; } break;                 \
    case 7: { const int DD = 7; __VA_ARGS__; } break;                 \
    case 8: { const int DD = 8; __VA_ARGS__; } break;                 \
    default: break;                                            \
  }

template <int D> static Ray<D> load_ray(const float* r) {
  Ray<D> ray;
  for (int i = 0; i < D; ++i) { ray.origin.v[i] = r[i]; ray.dir.v[i] = r[D + i]; }
  return ray;
}

template <int D> static void write_hit(const Hit<D>& h, bool have, uint64_t i, int64_t* leaf, float* t, float* point, float* normal, int64_t* sub) {
  if (leaf) leaf[i] = have ? h.leaf : -1;
  if (t) t[i] = have ? h.toi : INFINITY;
  for (int k = 0; k < D; ++k) {
    if (point) point[i * D + k] = have ? h.point.v[k] : 0.0f;
    if (normal) normal[i * D + k] = have ? h.normal.v[k] : 0.0f;
  }
  if (sub) sub[i] = have ? h.sub : -1;
}

extern "C" {

// --- SplitMix64 stream, as the reference tests draw it --------------------
void orc_rng_range(uint64_t seed, float lo, float hi, uint64_t n, float* out) {
  Rng r{seed};
  for (uint64_t i = 0; i < n; ++i) out[i] = r.range(lo, hi);
}
void orc_rng_u64(uint64_t seed, uint64_t n, uint64_t* out) {
  Rng r{seed};
  for (uint64_t i = 0; i < n; ++i) out[i] = r.next_u64();
}
// Ray::new normalisation (ray.rs:18-23) of n directions of dimension D, in place.
void orc_normalize(int D, float* v, uint64_t n) {
  ORC_DISPATCH(D, {
    for (uint64_t i = 0; i < n; ++i) {
      Point<DD> p; for (int k = 0; k < DD; ++k) p.v[k] = v[i * DD + k];
      p = normalize_or_zero(p);
      for (int k = 0; k < DD; ++k) v[i * DD + k] = p.v[k];
    }
  });
}

// --- geometry -------------------------------------------------------------
// Marks a geometry `guarded` (the engine's SQ_HITS_GUARDED) or plain (default).
void orc_geom_set_guarded(void* h, int guarded) {
  OGeom* g = (OGeom*)h;
  ORC_DISPATCH(g->D, { ((Geometry<DD>*)g->p)->guarded = guarded != 0; });
}
void* orc_geom_new(int D, int kind, const float* prims, uint32_t n) {
  OGeom* g = new OGeom{D, nullptr};
  ORC_DISPATCH(D, {
    Geometry<DD>* p = new Geometry<DD>();
    p->kind = kind; p->n = n;
    p->prims.assign(prims, prims + (size_t)n * p->stride());
    g->p = p;
  });
  return g;
}
// A consumer geometry: records [n][2D + pf] (world box, then the consumer's
// floats) tested by `fn` (an address, e.g. from a shared object the caller built).
void* orc_geom_new_consumer(int D, const float* records, uint32_t n, int pf, uint64_t fn) {
  OGeom* g = new OGeom{D, nullptr};
  ORC_DISPATCH(D, {
    Geometry<DD>* p = new Geometry<DD>();
    p->kind = KIND_CONSUMER; p->n = n; p->consumer_pf = pf;
    p->consumer = reinterpret_cast<ConsumerNarrow>(fn);
    p->prims.assign(records, records + (size_t)n * p->stride());
    g->p = p;
  });
  return g;
}
// Adds the consumer's shape-cast and overlap functions (addresses; 0 = none).
void orc_geom_set_consumer_cast_overlap(void* h, uint64_t cast_fn, uint64_t overlap_fn) {
  OGeom* g = (OGeom*)h;
  ORC_DISPATCH(g->D, {
    ((Geometry<DD>*)g->p)->consumer_cast = reinterpret_cast<ConsumerCast>(cast_fn);
    ((Geometry<DD>*)g->p)->consumer_ov = reinterpret_cast<ConsumerOverlap>(overlap_fn);
  });
}
// Adds the consumer's crossings function (an address; 0 = the nearest hit only).
void orc_geom_set_consumer_crossings(void* h, uint64_t fn) {
  OGeom* g = (OGeom*)h;
  ORC_DISPATCH(g->D, { ((Geometry<DD>*)g->p)->consumer_x = reinterpret_cast<ConsumerCrossings>(fn); });
}
void orc_geom_set_layers(void* h, const uint64_t* layers) {
  OGeom* g = (OGeom*)h;
  ORC_DISPATCH(g->D, { Geometry<DD>* p = (Geometry<DD>*)g->p; if (layers) p->layers.assign(layers, layers + p->n); else p->layers.clear(); });
}
void orc_geom_set_prims(void* h, const float* prims) {
  OGeom* g = (OGeom*)h;
  ORC_DISPATCH(g->D, { Geometry<DD>* p = (Geometry<DD>*)g->p; std::memcpy(p->prims.data(), prims, p->prims.size() * 4); });
}
void orc_geom_world_aabbs(void* h, float* out) {
  OGeom* g = (OGeom*)h;
  ORC_DISPATCH(g->D, {
    Geometry<DD>* p = (Geometry<DD>*)g->p;
    for (size_t i = 0; i < p->n; ++i) { Aabb<DD> b = p->world_aabb(i); for (int k = 0; k < DD; ++k) { out[i * 2 * DD + k] = b.min.v[k]; out[i * 2 * DD + DD + k] = b.max.v[k]; } }
  });
}
void orc_geom_free(void* h) {
  OGeom* g = (OGeom*)h;
  ORC_DISPATCH(g->D, { delete (Geometry<DD>*)g->p; });
  delete g;
}

// --- Bvh ------------------------------------------------------------------
void* orc_bvh_build(void* gh) {
  OGeom* g = (OGeom*)gh;
  OBvh* b = new OBvh{g->D, nullptr};
  ORC_DISPATCH(g->D, { Bvh<DD>* t = new Bvh<DD>(); t->build(*(Geometry<DD>*)g->p); b->p = t; });
  return b;
}
void orc_bvh_refit(void* bh, void* gh) {
  OBvh* b = (OBvh*)bh; OGeom* g = (OGeom*)gh;
  ORC_DISPATCH(b->D, { ((Bvh<DD>*)b->p)->refit(*(Geometry<DD>*)g->p); });
}
uint64_t orc_bvh_node_count(void* bh) {
  OBvh* b = (OBvh*)bh; uint64_t n = 0;
  ORC_DISPATCH(b->D, { n = ((Bvh<DD>*)b->p)->nodes.size(); });
  return n;
}
// bounds: [nn][2D] (min then max); child_a, child_b, prim_count: [nn]; indices: [n]
void orc_bvh_export(void* bh, float* bounds, uint32_t* ca, uint32_t* cb, uint32_t* pc, uint32_t* indices) {
  OBvh* b = (OBvh*)bh;
  ORC_DISPATCH(b->D, {
    Bvh<DD>* t = (Bvh<DD>*)b->p;
    for (size_t i = 0; i < t->nodes.size(); ++i) {
      const Node<DD>& nd = t->nodes[i];
      for (int k = 0; k < DD; ++k) { bounds[i * 2 * DD + k] = nd.bounds.min.v[k]; bounds[i * 2 * DD + DD + k] = nd.bounds.max.v[k]; }
      ca[i] = nd.child_a; cb[i] = nd.child_b; pc[i] = nd.prim_count;
    }
    std::memcpy(indices, t->prim_indices.data(), t->prim_indices.size() * 4);
  });
}
// TreeStats (stats.rs:15-32, bvh.rs:478-489): node, leaf, prim, depth, bytes; quality
void orc_bvh_stats(void* bh, uint64_t* out5, float* quality) {
  OBvh* b = (OBvh*)bh;
  ORC_DISPATCH(b->D, {
    Bvh<DD>* t = (Bvh<DD>*)b->p;
    out5[0] = t->nodes.size();
    tree_structural(t->nodes, &out5[1], &out5[2], &out5[3]);
    // size_of::<Node<D>>() = 8D + 12 rounded up to 4-byte alignment
    out5[4] = t->nodes.size() * (8 * DD + 12) + t->prim_indices.size() * 4;
    *quality = tree_interior_surface_sum(t->nodes);
  });
}
void orc_bvh_free(void* bh) {
  OBvh* b = (OBvh*)bh;
  ORC_DISPATCH(b->D, { delete (Bvh<DD>*)b->p; });
  delete b;
}

// --- ray queries ------------------------------------------------------------
// rays: [n][2D] (origin then dir).  Misses write leaf -1 and t = +inf.
// threads > 1 splits into contiguous chunks of ceil(n / workers), one thread
// each, as bvh.rs:160-193.  naive != 0 runs the O(N) scan instead.
void orc_raycast_nearest_batch(void* bh, void* gh, const float* rays, uint64_t n, float max_toi, uint64_t mask, int64_t exclude,
                               int threads, int naive, int64_t* out_leaf, float* out_t, float* out_point, float* out_normal,
                               int64_t* out_sub, uint64_t* out_stats) {
  OBvh* b = (OBvh*)bh; OGeom* g = (OGeom*)gh;
  Filter f{mask, exclude};
  ORC_DISPATCH(g->D, {
    const Bvh<DD>* t = b ? (const Bvh<DD>*)b->p : nullptr;
    const Geometry<DD>* geo = (const Geometry<DD>*)g->p;
    auto work = [&](uint64_t lo, uint64_t hi) {
      for (uint64_t i = lo; i < hi; ++i) {
        Ray<DD> ray = load_ray<DD>(rays + i * 2 * DD);
        Hit<DD> h; bool have;
        Stats st{0, 0, 0, 0};
        if (naive) have = naive_nearest(*geo, ray, max_toi, f, &h);
        else have = t->raycast_nearest(*geo, ray, max_toi, f, &h, out_stats ? &st : nullptr);
        write_hit(h, have, i, out_leaf, out_t, out_point, out_normal, out_sub);
        if (out_stats) { out_stats[4 * i] = st.nodes_visited; out_stats[4 * i + 1] = st.aabb_tests; out_stats[4 * i + 2] = st.narrow_tests; out_stats[4 * i + 3] = st.hits; }
      }
    };
    uint64_t workers = (uint64_t)std::max(1, threads);
    workers = std::min(workers, std::max<uint64_t>(n, 1));
    if (workers <= 1) work(0, n);
    else {
      uint64_t chunk = (n + workers - 1) / workers;
      std::vector<std::thread> th;
      for (uint64_t lo = 0; lo < n; lo += chunk) th.emplace_back(work, lo, std::min(n, lo + chunk));
      for (auto& x : th) x.join();
    }
  });
}

// All hits (crossings != 0: all crossings), sorted by (toi, leaf) per ray.
// Returns the total hit count; writes offsets[n+1] always, and the hit arrays
// only when cap >= total.
uint64_t orc_raycast_all_batch(void* bh, void* gh, const float* rays, uint64_t n, float max_toi, uint64_t mask, int64_t exclude,
                               int naive, int crossings, uint64_t* offsets, int64_t* out_leaf, float* out_t, uint64_t cap) {
  OBvh* b = (OBvh*)bh; OGeom* g = (OGeom*)gh;
  Filter f{mask, exclude};
  uint64_t total = 0;
  ORC_DISPATCH(g->D, {
    const Bvh<DD>* t = b ? (const Bvh<DD>*)b->p : nullptr;
    const Geometry<DD>* geo = (const Geometry<DD>*)g->p;
    std::vector<std::vector<Hit<DD>>> all(n);
    for (uint64_t i = 0; i < n; ++i) {
      Ray<DD> ray = load_ray<DD>(rays + i * 2 * DD);
      if (naive) naive_all(*geo, ray, max_toi, f, crossings != 0, all[i]);
      else t->raycast_all(*geo, ray, max_toi, f, crossings != 0, all[i], nullptr);
    }
    offsets[0] = 0;
    for (uint64_t i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + all[i].size();
    total = offsets[n];
    if (cap >= total)
      for (uint64_t i = 0; i < n; ++i)
        for (size_t k = 0; k < all[i].size(); ++k) { out_leaf[offsets[i] + k] = all[i][k].leaf; out_t[offsets[i] + k] = all[i][k].toi; }
  });
  return total;
}

// Bvh::raycast_all_stats' counters per ray (bvh.rs:206-215): out_stats[n][4].
void orc_raycast_all_stats(void* bh, void* gh, const float* rays, uint64_t n, float max_toi, uint64_t mask, int64_t exclude,
                           uint64_t* out_stats) {
  OBvh* b = (OBvh*)bh; OGeom* g = (OGeom*)gh;
  Filter f{mask, exclude};
  ORC_DISPATCH(g->D, {
    const Bvh<DD>* t = (const Bvh<DD>*)b->p;
    const Geometry<DD>* geo = (const Geometry<DD>*)g->p;
    for (uint64_t i = 0; i < n; ++i) {
      std::vector<Hit<DD>> hits;
      Stats st{0, 0, 0, 0};
      t->raycast_all(*geo, load_ray<DD>(rays + i * 2 * DD), max_toi, f, false, hits, &st);
      out_stats[4 * i] = st.nodes_visited; out_stats[4 * i + 1] = st.aabb_tests;
      out_stats[4 * i + 2] = st.narrow_tests; out_stats[4 * i + 3] = st.hits;
    }
  });
}

// regions: [n][2D] (min then max).  Leaves per query in ascending order.
uint64_t orc_overlap_batch(void* bh, void* gh, const float* regions, uint64_t n, uint64_t mask, int64_t exclude, int naive,
                           uint64_t* offsets, int64_t* out_leaf, uint64_t cap) {
  OBvh* b = (OBvh*)bh; OGeom* g = (OGeom*)gh;
  Filter f{mask, exclude};
  uint64_t total = 0;
  ORC_DISPATCH(g->D, {
    const Bvh<DD>* t = b ? (const Bvh<DD>*)b->p : nullptr;
    const Geometry<DD>* geo = (const Geometry<DD>*)g->p;
    std::vector<std::vector<int64_t>> all(n);
    for (uint64_t i = 0; i < n; ++i) {
      Aabb<DD> r; for (int k = 0; k < DD; ++k) { r.min.v[k] = regions[i * 2 * DD + k]; r.max.v[k] = regions[i * 2 * DD + DD + k]; }
      if (naive) naive_overlap(*geo, r, f, all[i]); else t->overlap(*geo, r, f, all[i]);
    }
    offsets[0] = 0;
    for (uint64_t i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + all[i].size();
    total = offsets[n];
    if (cap >= total)
      for (uint64_t i = 0; i < n; ++i) std::memcpy(out_leaf + offsets[i], all[i].data(), all[i].size() * 8);
  });
  return total;
}

// casts: [n][4D] (origin, dir, probe min, probe max); max_toi per query.
void orc_shapecast_nearest_batch(void* bh, void* gh, const float* casts, uint64_t n, const float* max_toi, uint64_t mask, int64_t exclude,
                                 int naive, int64_t* out_leaf, float* out_t, float* out_point, float* out_normal) {
  OBvh* b = (OBvh*)bh; OGeom* g = (OGeom*)gh;
  Filter f{mask, exclude};
  ORC_DISPATCH(g->D, {
    const Bvh<DD>* t = b ? (const Bvh<DD>*)b->p : nullptr;
    const Geometry<DD>* geo = (const Geometry<DD>*)g->p;
    for (uint64_t i = 0; i < n; ++i) {
      const float* c = casts + i * 4 * DD;
      ShapeCast<DD> sc;
      for (int k = 0; k < DD; ++k) { sc.origin.v[k] = c[k]; sc.dir.v[k] = c[DD + k]; sc.aabb.min.v[k] = c[2 * DD + k]; sc.aabb.max.v[k] = c[3 * DD + k]; }
      Hit<DD> h; bool have;
      if (naive) have = naive_shapecast_nearest(*geo, sc, max_toi[i], f, &h);
      else have = t->shapecast_nearest(*geo, sc, max_toi[i], f, &h);
      write_hit(h, have, i, out_leaf, out_t, out_point, out_normal, (int64_t*)nullptr);
    }
  });
}

uint64_t orc_shapecast_all_batch(void* bh, void* gh, const float* casts, uint64_t n, const float* max_toi, uint64_t mask, int64_t exclude,
                                 int naive, uint64_t* offsets, int64_t* out_leaf, float* out_t, uint64_t cap) {
  OBvh* b = (OBvh*)bh; OGeom* g = (OGeom*)gh;
  Filter f{mask, exclude};
  uint64_t total = 0;
  ORC_DISPATCH(g->D, {
    const Bvh<DD>* t = b ? (const Bvh<DD>*)b->p : nullptr;
    const Geometry<DD>* geo = (const Geometry<DD>*)g->p;
    std::vector<std::vector<Hit<DD>>> all(n);
    for (uint64_t i = 0; i < n; ++i) {
      const float* c = casts + i * 4 * DD;
      ShapeCast<DD> sc;
      for (int k = 0; k < DD; ++k) { sc.origin.v[k] = c[k]; sc.dir.v[k] = c[DD + k]; sc.aabb.min.v[k] = c[2 * DD + k]; sc.aabb.max.v[k] = c[3 * DD + k]; }
      if (naive) naive_shapecast_all(*geo, sc, max_toi[i], f, all[i]); else t->shapecast_all(*geo, sc, max_toi[i], f, all[i]);
    }
    offsets[0] = 0;
    for (uint64_t i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + all[i].size();
    total = offsets[n];
    if (cap >= total)
      for (uint64_t i = 0; i < n; ++i)
        for (size_t k = 0; k < all[i].size(); ++k) { out_leaf[offsets[i] + k] = all[i][k].leaf; out_t[offsets[i] + k] = all[i][k].toi; }
  });
  return total;
}

// Order-independent minimal work per ray at limit t_limit[i] (t* or max_toi).
void orc_required_work(void* bh, const float* rays, uint64_t n, const float* t_limit, uint64_t* out_nodes, uint64_t* out_prims, int threads) {
  OBvh* b = (OBvh*)bh;
  ORC_DISPATCH(b->D, {
    const Bvh<DD>* t = (const Bvh<DD>*)b->p;
    auto work = [&](uint64_t lo, uint64_t hi) {
      for (uint64_t i = lo; i < hi; ++i) t->required_work(load_ray<DD>(rays + i * 2 * DD), t_limit[i], &out_nodes[i], &out_prims[i]);
    };
    uint64_t workers = (uint64_t)std::max(1, threads);
    workers = std::min(workers, std::max<uint64_t>(n, 1));
    uint64_t chunk = (n + workers - 1) / workers;
    std::vector<std::thread> th;
    for (uint64_t lo = 0; lo < n; lo += chunk) th.emplace_back(work, lo, std::min(n, lo + chunk));
    for (auto& x : th) x.join();
  });
}

// --- instances and Tlas -----------------------------------------------------
// rot: [n][D][D] row-major; trans: [n][D].
// INST_BOX: halves [n][D].  INST_BALL: radii [n] in halves.  INST_MESH (D = 3): mesh_tri_off [n_meshes + 1]
// into tris [T][9]; mesh_of [n]; scale [n].
void* orc_inst_new(int D, int kind, uint64_t n, const float* rot, const float* trans, const float* halves,
                   uint32_t n_meshes, const uint64_t* mesh_tri_off, const float* tris, const uint32_t* mesh_of, const float* scale) {
  OInst* o = new OInst{D, nullptr, nullptr};
  ORC_DISPATCH(D, {
    Instances<DD>* s = new Instances<DD>();
    s->kind = kind == INST_BOX_GUARDED ? INST_BOX : kind;
    s->guarded = kind == INST_BOX_GUARDED;
    s->xf.resize(n);
    for (uint64_t i = 0; i < n; ++i) {
      for (int a = 0; a < DD; ++a) for (int c = 0; c < DD; ++c) s->xf[i].rot[a][c] = rot[i * DD * DD + a * DD + c];
      for (int a = 0; a < DD; ++a) s->xf[i].trans.v[a] = trans[i * DD + a];
    }
    if (s->is_box()) {
      s->halves.resize(n);
      for (uint64_t i = 0; i < n; ++i) for (int a = 0; a < DD; ++a) s->halves[i].v[a] = halves[i * DD + a];
    } else if (kind == INST_BALL) {
      s->radius.assign(halves, halves + n);
    } else {
      for (uint32_t m = 0; m < n_meshes; ++m) {
        Geometry<DD>* g = new Geometry<DD>();
        g->kind = KIND_TRI; g->n = (uint32_t)(mesh_tri_off[m + 1] - mesh_tri_off[m]);
        g->prims.assign(tris + mesh_tri_off[m] * 9, tris + mesh_tri_off[m + 1] * 9);
        Bvh<DD>* bv = new Bvh<DD>(); bv->build(*g);
        Aabb<DD> mb = Aabb<DD>::empty();
        for (size_t k = 0; k < g->n; ++k) mb = mb.unite(g->world_aabb(k));
        s->meshes.emplace_back(g); s->mesh_bvh.emplace_back(bv); s->mesh_box.push_back(mb);
      }
      s->mesh_of.assign(mesh_of, mesh_of + n);
      s->scale.assign(scale, scale + n);
    }
    o->p = s;
    Tlas<DD>* t = new Tlas<DD>(); t->build(*s); o->tlas = t;
  });
  return o;
}
// Marks an instance set `guarded` (the engine's SQ_HITS_GUARDED: instance hits
// kept inside their world boxes, mesh triangles guarded) or plain (default).
void orc_inst_set_guarded(void* h, int guarded) {
  OInst* o = (OInst*)h;
  ORC_DISPATCH(o->D, {
    Instances<DD>* s = (Instances<DD>*)o->p;
    s->guarded = guarded != 0;
    for (auto& m : s->meshes) m->guarded = guarded != 0;
  });
}
void orc_inst_set_transforms(void* h, const float* rot, const float* trans, const float* halves) {
  OInst* o = (OInst*)h;
  ORC_DISPATCH(o->D, {
    Instances<DD>* s = (Instances<DD>*)o->p;
    for (size_t i = 0; i < s->xf.size(); ++i) {
      for (int a = 0; a < DD; ++a) for (int c = 0; c < DD; ++c) s->xf[i].rot[a][c] = rot[i * DD * DD + a * DD + c];
      for (int a = 0; a < DD; ++a) s->xf[i].trans.v[a] = trans[i * DD + a];
      if (halves && s->is_box()) for (int a = 0; a < DD; ++a) s->halves[i].v[a] = halves[i * DD + a];
      if (halves && s->kind == INST_BALL) s->radius[i] = halves[i];
    }
  });
}
void orc_tlas_refit(void* h) {
  OInst* o = (OInst*)h;
  ORC_DISPATCH(o->D, { ((Tlas<DD>*)o->tlas)->refit(*(Instances<DD>*)o->p); });
}
void orc_tlas_rebuild(void* h) {
  OInst* o = (OInst*)h;
  ORC_DISPATCH(o->D, { ((Tlas<DD>*)o->tlas)->build(*(Instances<DD>*)o->p); });
}
int orc_tlas_needs_rebuild(void* h, float factor) {
  OInst* o = (OInst*)h; int r = 0;
  ORC_DISPATCH(o->D, { r = ((Tlas<DD>*)o->tlas)->needs_rebuild(factor) ? 1 : 0; });
  return r;
}
uint64_t orc_tlas_node_count(void* h) {
  OInst* o = (OInst*)h; uint64_t n = 0;
  ORC_DISPATCH(o->D, { n = ((Tlas<DD>*)o->tlas)->nodes.size(); });
  return n;
}
void orc_tlas_export(void* h, float* bounds, uint32_t* ca, uint32_t* cb, uint32_t* pc, uint32_t* indices, float* world) {
  OInst* o = (OInst*)h;
  ORC_DISPATCH(o->D, {
    Tlas<DD>* t = (Tlas<DD>*)o->tlas;
    for (size_t i = 0; i < t->nodes.size(); ++i) {
      const Node<DD>& nd = t->nodes[i];
      for (int k = 0; k < DD; ++k) { bounds[i * 2 * DD + k] = nd.bounds.min.v[k]; bounds[i * 2 * DD + DD + k] = nd.bounds.max.v[k]; }
      ca[i] = nd.child_a; cb[i] = nd.child_b; pc[i] = nd.prim_count;
    }
    std::memcpy(indices, t->idx.data(), t->idx.size() * 4);
    if (world) for (size_t i = 0; i < t->world.size(); ++i) for (int k = 0; k < DD; ++k) { world[i * 2 * DD + k] = t->world[i].min.v[k]; world[i * 2 * DD + DD + k] = t->world[i].max.v[k]; }
  });
}
void orc_tlas_raycast_nearest_batch(void* h, const float* rays, uint64_t n, float max_toi, int naive, int threads,
                                    int64_t* out_leaf, float* out_t, float* out_point, float* out_normal, int64_t* out_sub) {
  OInst* o = (OInst*)h;
  Filter f{~0ull, -1};
  ORC_DISPATCH(o->D, {
    const Tlas<DD>* t = (const Tlas<DD>*)o->tlas;
    const Instances<DD>* s = (const Instances<DD>*)o->p;
    auto work = [&](uint64_t lo, uint64_t hi) {
      for (uint64_t i = lo; i < hi; ++i) {
        Ray<DD> ray = load_ray<DD>(rays + i * 2 * DD);
        Hit<DD> hh; bool have = naive ? naive_tlas_nearest(*t, *s, ray, max_toi, &hh) : t->raycast_nearest(*s, ray, max_toi, f, &hh);
        write_hit(hh, have, i, out_leaf, out_t, out_point, out_normal, out_sub);
      }
    };
    uint64_t workers = (uint64_t)std::max(1, threads);
    workers = std::min(workers, std::max<uint64_t>(n, 1));
    uint64_t chunk = (n + workers - 1) / workers;
    std::vector<std::thread> th;
    for (uint64_t lo = 0; lo < n; lo += chunk) th.emplace_back(work, lo, std::min(n, lo + chunk));
    for (auto& x : th) x.join();
  });
}
uint64_t orc_tlas_raycast_all_batch(void* h, const float* rays, uint64_t n, float max_toi, uint64_t* offsets, int64_t* out_leaf,
                                    float* out_t, int64_t* out_sub, uint64_t cap) {
  OInst* o = (OInst*)h;
  Filter f{~0ull, -1};
  uint64_t total = 0;
  ORC_DISPATCH(o->D, {
    const Tlas<DD>* t = (const Tlas<DD>*)o->tlas;
    const Instances<DD>* s = (const Instances<DD>*)o->p;
    std::vector<std::vector<Hit<DD>>> all(n);
    for (uint64_t i = 0; i < n; ++i) t->raycast_all(*s, load_ray<DD>(rays + i * 2 * DD), max_toi, f, all[i]);
    offsets[0] = 0;
    for (uint64_t i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + all[i].size();
    total = offsets[n];
    if (cap >= total)
      for (uint64_t i = 0; i < n; ++i)
        for (size_t k = 0; k < all[i].size(); ++k) {
          out_leaf[offsets[i] + k] = all[i][k].leaf; out_t[offsets[i] + k] = all[i][k].toi;
          if (out_sub) out_sub[offsets[i] + k] = all[i][k].sub;
        }
  });
  return total;
}
uint64_t orc_tlas_overlap_batch(void* h, const float* regions, uint64_t n, uint64_t* offsets, int64_t* out_leaf, uint64_t cap) {
  OInst* o = (OInst*)h;
  Filter f{~0ull, -1};
  uint64_t total = 0;
  ORC_DISPATCH(o->D, {
    const Tlas<DD>* t = (const Tlas<DD>*)o->tlas;
    const Instances<DD>* s = (const Instances<DD>*)o->p;
    std::vector<std::vector<int64_t>> all(n);
    for (uint64_t i = 0; i < n; ++i) {
      Aabb<DD> r; for (int k = 0; k < DD; ++k) { r.min.v[k] = regions[i * 2 * DD + k]; r.max.v[k] = regions[i * 2 * DD + DD + k]; }
      t->overlap(*s, r, f, all[i]);
    }
    offsets[0] = 0;
    for (uint64_t i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + all[i].size();
    total = offsets[n];
    if (cap >= total) for (uint64_t i = 0; i < n; ++i) std::memcpy(out_leaf + offsets[i], all[i].data(), all[i].size() * 8);
  });
  return total;
}
// Tlas::shapecast_nearest / shapecast_all over casts [n][4D] (origin, dir,
// probe box min, max); all = 0: nearest (leaf -1 on a miss), else CSR.
uint64_t orc_tlas_shapecast_batch(void* h, const float* casts, uint64_t n, float max_toi, int all, uint64_t* offsets,
                                  int64_t* out_leaf, float* out_t, uint64_t cap) {
  OInst* o = (OInst*)h;
  Filter f{~0ull, -1};
  uint64_t total = 0;
  ORC_DISPATCH(o->D, {
    const Tlas<DD>* t = (const Tlas<DD>*)o->tlas;
    const Instances<DD>* s = (const Instances<DD>*)o->p;
    auto cast_of = [&](uint64_t i) {
      ShapeCast<DD> c;
      const float* q = casts + i * 4 * DD;
      for (int k = 0; k < DD; ++k) {
        c.origin.v[k] = q[k]; c.dir.v[k] = q[DD + k]; c.aabb.min.v[k] = q[2 * DD + k]; c.aabb.max.v[k] = q[3 * DD + k];
      }
      return c;
    };
    if (!all) {
      for (uint64_t i = 0; i < n; ++i) {
        Hit<DD> hh;
        if (t->shapecast_nearest(*s, cast_of(i), max_toi, f, &hh)) { out_leaf[i] = hh.leaf; out_t[i] = hh.toi; }
        else { out_leaf[i] = -1; out_t[i] = INFINITY; }
      }
      total = n;
    } else {
      std::vector<std::vector<Hit<DD>>> hits(n);
      for (uint64_t i = 0; i < n; ++i) t->shapecast_all(*s, cast_of(i), max_toi, f, hits[i]);
      offsets[0] = 0;
      for (uint64_t i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + hits[i].size();
      total = offsets[n];
      if (cap >= total)
        for (uint64_t i = 0; i < n; ++i)
          for (size_t k = 0; k < hits[i].size(); ++k) { out_leaf[offsets[i] + k] = hits[i][k].leaf; out_t[offsets[i] + k] = hits[i][k].toi; }
    }
  });
  return total;
}
void orc_inst_free(void* h) {
  OInst* o = (OInst*)h;
  ORC_DISPATCH(o->D, { delete (Instances<DD>*)o->p; delete (Tlas<DD>*)o->tlas; });
  delete o;
}

// Isometry helpers for closed-form tests (isometry.rs:56-97).
void orc_iso_apply(int D, const float* rot, const float* trans, const float* p, int op, float* out) {
  ORC_DISPATCH(D, {
    Isometry<DD> iso;
    for (int a = 0; a < DD; ++a) { for (int c = 0; c < DD; ++c) iso.rot[a][c] = rot[a * DD + c]; iso.trans.v[a] = trans[a]; }
    Point<DD> x; for (int a = 0; a < DD; ++a) x.v[a] = p[a];
    Point<DD> y;
    if (op == 0) y = iso.transform_point(x);
    else if (op == 1) y = iso.transform_vector(x);
    else if (op == 2) y = iso.inverse_transform_point(x);
    else y = iso.inverse_transform_vector(x);
    for (int a = 0; a < DD; ++a) out[a] = y.v[a];
  });
}
void orc_iso_transform_aabb(int D, const float* rot, const float* trans, const float* box, float* out) {
  ORC_DISPATCH(D, {
    Isometry<DD> iso;
    for (int a = 0; a < DD; ++a) { for (int c = 0; c < DD; ++c) iso.rot[a][c] = rot[a * DD + c]; iso.trans.v[a] = trans[a]; }
    Aabb<DD> b; for (int a = 0; a < DD; ++a) { b.min.v[a] = box[a]; b.max.v[a] = box[DD + a]; }
    Aabb<DD> w = iso.transform_aabb(b);
    for (int a = 0; a < DD; ++a) { out[a] = w.min.v[a]; out[DD + a] = w.max.v[a]; }
  });
}
// Aabb::ray_intersect for one box and one ray (returns 1 on hit).
int orc_ray_intersect(int D, const float* box, const float* ray, float max_toi, float* tmin, float* tmax) {
  int hit = 0;
  ORC_DISPATCH(D, {
    Aabb<DD> b; for (int a = 0; a < DD; ++a) { b.min.v[a] = box[a]; b.max.v[a] = box[DD + a]; }
    hit = b.ray_intersect(load_ray<DD>(ray), max_toi, tmin, tmax) ? 1 : 0;
  });
  return hit;
}

}  // extern "C"
