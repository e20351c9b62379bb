// traverse.h -- per-ray query logic of the engine, shared by the HIP kernels
// (engine.hip) and a CPU build of the same source used by the tests.
//
// FP32 arithmetic follows the reference operation for operation (no
// contraction: every TU including this header is built with
// -ffp-contract=off; divisions and square roots are correctly rounded):
//   slab test            crates/spatial-query/src/maths/aabb.rs:118-148
//   ball narrow phase    crates/spatial-query/tests/differential.rs:61-82
//   Moeller-Trumbore     crates/viewport-lib-query/src/lib.rs:279-306
//   nearest update       crates/spatial-query/src/accel/bvh.rs:106-115 with
//                        hit_ordering, query/hit.rs:77-84
//
// Two walk orders.  ORDER_REF walks the reference's own order: push child a,
// then b, pop LIFO (so b's subtree first), a popped entry skipped when it lies
// beyond the current best; with the narrow phases unguarded this is
// bvh.rs:73-124 step for step and returns the reference's answer for any
// provider, including hits that fall outside their own AABB.  ORDER_NEAR
// descends near-first; it returns the same answer when every accepted hit lies
// inside its own AABB's slab interval (AABB hits are that interval's entry;
// with GUARD, SPHERE and TRIANGLE hits are kept only inside it), since such a
// hit lies inside every ancestor box and pruning by "entry <= current best
// toi" can then never drop a hit that beats the current best under
// (toi, leaf): the result is the minimum over all accepted hits, whatever the
// visiting order.  Queries whose limit never shrinks (any-hit, all hits,
// crossings, overlap) visit the same set in any order and run unguarded.
#pragma once

#include <cstdint>
#if !defined(__HIPCC_RTC__)
// (hipRTC, which compiles this header into consumer narrow-phase kernels,
// provides the HIP runtime itself and no <cstring>)
#include <cstring>
#endif

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
#if !defined(__HIPCC_RTC__)
#include <hip/hip_runtime.h>
#endif
#define SQ_HD __host__ __device__ __forceinline__
// IEEE division and square root (correctly rounded): plain `/` and
// __builtin_sqrtf lower to the correctly rounded sequences under HIP's default
// -fhip-fp32-correctly-rounded-divide-sqrt.  (__fsqrt_rn is the native,
// approximate sqrt in this toolchain unless OCML_BASIC_ROUNDED_OPERATIONS.)
#define SQ_DIV_RN(a, b) ((a) / (b))
#define SQ_SQRT_RN(a) __builtin_sqrtf(a)
#else
#include <cmath>
#include <cstring>
#define SQ_HD inline
#define SQ_DIV_RN(a, b) ((a) / (b))
#define SQ_SQRT_RN(a) sqrtf(a)
#endif

namespace sqd {

constexpr float F32_EPSILON = 1.1920929e-7f;   // f32::EPSILON
constexpr float SQ_INF = __builtin_huge_valf();
constexpr uint32_t LEAF_BIT = 0x80000000u;
constexpr uint32_t FIRST_MASK = (1u << 27) - 1;
constexpr uint32_t NO_LEAF = 0xFFFFFFFFu;

enum { KIND_AABB = 0, KIND_SPHERE = 1, KIND_TRI = 2 };

SQ_HD uint32_t fbits(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __float_as_uint(f);
#else
  uint32_t u; std::memcpy(&u, &f, 4); return u;
#endif
}
SQ_HD float bitsf(uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __uint_as_float(u);
#else
  float f; std::memcpy(&f, &u, 4); return f;
#endif
}
// Scalar-generic pieces of the arithmetic below.  The FP32 walks instantiate
// it on float; the f64 walk (f64.hip) on double, as the reference's `f64`
// feature switches its Scalar (crates/spatial-query/src/maths/point.rs:14-18).
template <class R> struct Eps;
template <> struct Eps<float> { static constexpr float v = 1.1920929e-7f; };            // f32::EPSILON
template <> struct Eps<double> { static constexpr double v = 2.220446049250313e-16; };  // f64::EPSILON
SQ_HD float sq_abs(float x) { return fabsf(x); }
SQ_HD double sq_abs(double x) { return fabs(x); }
SQ_HD float sq_min(float a, float b) { return fminf(a, b); }
SQ_HD double sq_min(double a, double b) { return fmin(a, b); }
SQ_HD float sq_max(float a, float b) { return fmaxf(a, b); }
SQ_HD double sq_max(double a, double b) { return fmax(a, b); }
SQ_HD float sq_sqrt(float x) { return SQ_SQRT_RN(x); }
#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
SQ_HD double sq_sqrt(double x) { return __builtin_sqrt(x); }   // correctly rounded f64 sqrt
#else
SQ_HD double sq_sqrt(double x) { return std::sqrt(x); }
#endif

// f32::total_cmp as a signed key.
SQ_HD int32_t total_key(float f) {
  int32_t i = (int32_t)fbits(f);
  return i ^ (int32_t)(((uint32_t)(i >> 31)) >> 1);
}
// (toi, leaf) strictly before (bt, bl) under hit_ordering.
SQ_HD bool hit_before(float t, uint32_t leaf, float bt, uint32_t bl) {
  int32_t a = total_key(t), b = total_key(bt);
  return a < b || (a == b && leaf < bl);
}

// Outputs of nearest-hit queries (device arrays; point, normal, sub, any optional).
struct NearestOut {
  int32_t* leaf;
  float* toi;
  float* point;
  float* normal;
  int32_t* sub;
  uint8_t* any;
};

struct Filter {
  const uint64_t* layers;   // by original leaf id, or null (all ones)
  uint64_t mask;
  int64_t exclude;          // -1 = None
};

SQ_HD bool accepts(const Filter& f, uint32_t leaf) {
  if (f.exclude >= 0 && (int64_t)leaf == f.exclude) return false;
  uint64_t lay = f.layers ? f.layers[leaf] : ~0ull;
  return (lay & f.mask) != 0;
}

// Fetch n16 16-byte chunks (pair nodes and triangle records are 16-byte
// aligned) into registers: dwordx4 loads on the device.
template <int N16> SQ_HD void fetch16(const float* src, float* dst) {
#if defined(__HIP_DEVICE_COMPILE__)
  const float4* p = reinterpret_cast<const float4*>(src);
#pragma unroll
  for (int i = 0; i < N16; ++i) {
    const float4 v = p[i];
    dst[4 * i] = v.x; dst[4 * i + 1] = v.y; dst[4 * i + 2] = v.z; dst[4 * i + 3] = v.w;
  }
#else
  for (int i = 0; i < 4 * N16; ++i) dst[i] = src[i];
#endif
}

template <int D, class R = float> struct RayD {
  R o[D], d[D], inv[D];
  uint32_t par;             // bit k: |d[k]| < EPSILON (the slab's parallel branch)
};

template <int D, class R = float> SQ_HD void ray_setup(RayD<D, R>& r, const R* src) {
  r.par = 0;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    r.o[k] = src[k];
    r.d[k] = src[D + k];
    // The reference computes 1.0 / d per node; hoisting it gives the same bits.
    r.inv[k] = SQ_DIV_RN(R(1), r.d[k]);
    if (sq_abs(r.d[k]) < Eps<R>::v) r.par |= 1u << k;
  }
}

// ---------------------------------------------------------------------------
// Slab tests
// ---------------------------------------------------------------------------

// Exact reference slab (aabb.rs:118-148) on one box given as min[D], max[D]
// with stride `s` between the min and max of an axis.  Returns hit; writes the
// entry (t_min, exactly as the reference: +0 unless an axis raised it), the
// exit and the entry axis (-1 when no axis raised t_min).
template <int D, class R = float> SQ_HD bool slab_exact(const RayD<D, R>& r, const R* mn, const R* mx, R max_toi,
                                                    R& t_min_out, R& t_max_out, int& axis_out) {
  R t_min = R(0), t_max = max_toi;
  int axis = -1;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    if (r.par & (1u << k)) {
      if (r.o[k] < mn[k] || r.o[k] > mx[k]) return false;
    } else {
      R t1 = (mn[k] - r.o[k]) * r.inv[k];
      R t2 = (mx[k] - r.o[k]) * r.inv[k];
      if (t1 > t2) { R tmp = t1; t1 = t2; t2 = tmp; }
      if (t1 > t_min) { t_min = t1; axis = k; }
      if (t2 < t_max) t_max = t2;
      if (t_min > t_max) return false;
    }
  }
  t_min_out = t_min; t_max_out = t_max; axis_out = axis;
  return true;
}

// Pruning slab of both children of a pair node.  Pass/fail equals the exact
// slab's for finite inputs (min/max replace the strict compare-and-assign; the
// sign of a zero entry cannot change a pass/fail).  ta/tb are the entries.
template <int D> SQ_HD void slab_pair(const RayD<D>& r, const float* nd, float limit, bool& ha, bool& hb, float& ta, float& tb) {
  float amin = 0.0f, amax = limit, bmin = 0.0f, bmax = limit;
  if (r.par == 0) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const float o = r.o[k], iv = r.inv[k];
      float a1 = (nd[4 * k + 0] - o) * iv, a2 = (nd[4 * k + 1] - o) * iv;
      float b1 = (nd[4 * k + 2] - o) * iv, b2 = (nd[4 * k + 3] - o) * iv;
      amin = fmaxf(amin, fminf(a1, a2)); amax = fminf(amax, fmaxf(a1, a2));
      bmin = fmaxf(bmin, fminf(b1, b2)); bmax = fminf(bmax, fmaxf(b1, b2));
    }
  } else {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const float o = r.o[k];
      if (r.par & (1u << k)) {
        if (o < nd[4 * k + 0] || o > nd[4 * k + 1]) amin = SQ_INF;
        if (o < nd[4 * k + 2] || o > nd[4 * k + 3]) bmin = SQ_INF;
      } else {
        const float iv = r.inv[k];
        float a1 = (nd[4 * k + 0] - o) * iv, a2 = (nd[4 * k + 1] - o) * iv;
        float b1 = (nd[4 * k + 2] - o) * iv, b2 = (nd[4 * k + 3] - o) * iv;
        amin = fmaxf(amin, fminf(a1, a2)); amax = fminf(amax, fmaxf(a1, a2));
        bmin = fmaxf(bmin, fminf(b1, b2)); bmax = fminf(bmax, fmaxf(b1, b2));
      }
    }
  }
  ha = amin <= amax; hb = bmin <= bmax; ta = amin; tb = bmin;
}

// ---------------------------------------------------------------------------
// Narrow phases.  `p` points at the leaf-ordered device record (host_tree.h).
// Each returns true with toi when the primitive is hit within `limit`.
// ---------------------------------------------------------------------------

template <int D, class R = float> SQ_HD bool narrow_aabb(const RayD<D, R>& r, const R* p, R limit, R& t, int& axis) {
  R te, tx;
  if (!slab_exact<D>(r, p, p + D, limit, te, tx, axis)) return false;
  t = te;
  return true;
}

template <int D, class R = float> SQ_HD R dotD(const R* a, const R* b) {
  R acc = R(0);
#pragma unroll
  for (int k = 0; k < D; ++k) acc = acc + a[k] * b[k];
  return acc;
}

// Ball (differential.rs:61-82); GUARD keeps a hit only inside the slab
// interval of the ball's AABB.
template <int D, class R = float, bool GUARD = false> SQ_HD bool narrow_sphere(const RayD<D, R>& r, const R* p, R limit, R& t_out) {
  R oc[D];
#pragma unroll
  for (int k = 0; k < D; ++k) oc[k] = r.o[k] - p[k];
  const R rad = p[D];
  R b = dotD<D>(oc, r.d);
  R cc = dotD<D>(oc, oc) - rad * rad;
  R disc = b * b - cc;
  if (disc < R(0)) return false;
  R sq = sq_sqrt(disc);
  R t = -b - sq;
  if (t < R(0)) t = -b + sq;
  if (t < R(0) || t > limit) return false;
  if constexpr (GUARD) {
    // toi inside the slab interval of the ball's AABB (c - r, c + r)
    R mn[D], mx[D];
#pragma unroll
    for (int k = 0; k < D; ++k) { mn[k] = p[k] - rad; mx[k] = p[k] + rad; }
    R te, tx; int ax;
    if (!slab_exact<D>(r, mn, mx, limit, te, tx, ax)) return false;
    if (!(te <= t && t <= tx)) return false;
  }
  t_out = t;
  return true;
}

// Moeller-Trumbore (viewport-lib-query/src/lib.rs:281-306) with the reference
// core's `t > max_toi` reject; GUARD keeps a hit only inside the slab interval
// of the triangle's AABB.  p = v0.xyz,id, v1.xyz,_, v2.xyz,_
template <bool GUARD = false> SQ_HD bool narrow_tri(const RayD<3>& r, const float* p, float limit, float& t_out) {
  const float v0x = p[0], v0y = p[1], v0z = p[2];
  const float v1x = p[4], v1y = p[5], v1z = p[6];
  const float v2x = p[8], v2y = p[9], v2z = p[10];
  const float e1x = v1x - v0x, e1y = v1y - v0y, e1z = v1z - v0z;
  const float e2x = v2x - v0x, e2y = v2y - v0y, e2z = v2z - v0z;
  const float dx = r.d[0], dy = r.d[1], dz = r.d[2];
  // p = d x e2
  const float px = dy * e2z - dz * e2y;
  const float py = dz * e2x - dx * e2z;
  const float pz = dx * e2y - dy * e2x;
  const float det = ((0.0f + e1x * px) + e1y * py) + e1z * pz;
  if (fabsf(det) < 1e-8f) return false;
  const float inv = SQ_DIV_RN(1.0f, det);
  const float tx_ = r.o[0] - v0x, ty_ = r.o[1] - v0y, tz_ = r.o[2] - v0z;
  const float u = (((0.0f + tx_ * px) + ty_ * py) + tz_ * pz) * inv;
  if (!(u >= 0.0f && u <= 1.0f)) return false;
  // q = tvec x e1
  const float qx = ty_ * e1z - tz_ * e1y;
  const float qy = tz_ * e1x - tx_ * e1z;
  const float qz = tx_ * e1y - ty_ * e1x;
  const float v = (((0.0f + dx * qx) + dy * qy) + dz * qz) * inv;
  if (v < 0.0f || u + v > 1.0f) return false;
  const float t = (((0.0f + e2x * qx) + e2y * qy) + e2z * qz) * inv;
  if (!(t > 1e-6f)) return false;
  if (t > limit) return false;
  if constexpr (GUARD) {
    // toi inside the slab interval of the triangle's AABB
    float mn[3], mx[3];
    mn[0] = fminf(fminf(v0x, v1x), v2x); mx[0] = fmaxf(fmaxf(v0x, v1x), v2x);
    mn[1] = fminf(fminf(v0y, v1y), v2y); mx[1] = fmaxf(fmaxf(v0y, v1y), v2y);
    mn[2] = fminf(fminf(v0z, v1z), v2z); mx[2] = fmaxf(fmaxf(v0z, v1z), v2z);
    float te, tx; int ax;
    if (!slab_exact<3>(r, mn, mx, limit, te, tx, ax)) return false;
    if (!(te <= t && t <= tx)) return false;
  }
  t_out = t;
  return true;
}

template <int D, int KIND> struct Prim;
template <int D> struct Prim<D, KIND_AABB> { static constexpr int STRIDE = 2 * D; };
template <int D> struct Prim<D, KIND_SPHERE> { static constexpr int STRIDE = D + 1; };
template <> struct Prim<3, KIND_TRI> { static constexpr int STRIDE = 12; };

template <int D, int KIND, class R = float, bool GUARD = false> SQ_HD bool narrow(const RayD<D, R>& r, const R* p, R limit, R& t) {
  if constexpr (KIND == KIND_AABB) { int ax; return narrow_aabb<D>(r, p, limit, t, ax); }
  else if constexpr (KIND == KIND_SPHERE) return narrow_sphere<D, R, GUARD>(r, p, limit, t);
  else return narrow_tri<GUARD>(r, p, limit, t);
}

// True when the leaf-ordered record is a synthetic pair node: one part of a
// reference leaf split into a chain (host_tree.h), not a node of the reference.
template <int D> SQ_HD bool synthetic_pair(const float* nd) { return fbits(nd[4 * D + 3]) != 0u; }

// Normal and sub-object of the final hit (computed once per ray).
// AABB: entry-face normal (-sign(d) on the entry axis, zero when the origin is
// inside); SPHERE: normalize_or_zero(at(t) - c) (point.rs:82-89); TRIANGLE:
// normalize_or_zero(e1 x e2) turned against the ray.
template <int D, int KIND, class R = float> SQ_HD void hit_normal(const RayD<D, R>& r, const R* p, R t, R limit, R* n) {
#pragma unroll
  for (int k = 0; k < D; ++k) n[k] = R(0);
  if constexpr (KIND == KIND_AABB) {
    R tt; int ax = -1;
    narrow_aabb<D>(r, p, limit, tt, ax);
    // n[ax] = -sign(d[ax]) written as a select per axis (no dynamic indexing)
#pragma unroll
    for (int k = 0; k < D; ++k) n[k] = (k == ax) ? (r.d[k] > R(0) ? R(-1) : R(1)) : R(0);
  } else if constexpr (KIND == KIND_SPHERE) {
    R v[D];
#pragma unroll
    for (int k = 0; k < D; ++k) v[k] = (r.o[k] + r.d[k] * t) - p[k];
    R len = sq_sqrt(dotD<D>(v, v));
    if (len > Eps<R>::v) {
      R s = SQ_DIV_RN(R(1), len);
#pragma unroll
      for (int k = 0; k < D; ++k) n[k] = v[k] * s;
    }
  } else {
    const float e1x = p[4] - p[0], e1y = p[5] - p[1], e1z = p[6] - p[2];
    const float e2x = p[8] - p[0], e2y = p[9] - p[1], e2z = p[10] - p[2];
    float c[3] = {e1y * e2z - e1z * e2y, e1z * e2x - e1x * e2z, e1x * e2y - e1y * e2x};
    float len = SQ_SQRT_RN(dotD<3>(c, c));
    if (len > F32_EPSILON) {
      float s = SQ_DIV_RN(1.0f, len);
      for (int k = 0; k < 3; ++k) n[k] = c[k] * s;
      if (dotD<3>(n, r.d) > 0.0f) for (int k = 0; k < 3; ++k) n[k] = n[k] * -1.0f;
    }
  }
}

// ---------------------------------------------------------------------------
// Query policies: what a node test and a leaf test mean for each query kind.
// ---------------------------------------------------------------------------

// The AABB of one leaf-ordered primitive record (QueryGeometry::world_aabb).
template <int D, int KIND, class R = float> SQ_HD void prim_box(const R* p, R* mn, R* mx) {
  if constexpr (KIND == KIND_AABB) {
#pragma unroll
    for (int k = 0; k < D; ++k) { mn[k] = p[k]; mx[k] = p[D + k]; }
  } else if constexpr (KIND == KIND_SPHERE) {
#pragma unroll
    for (int k = 0; k < D; ++k) { mn[k] = p[k] - p[D]; mx[k] = p[k] + p[D]; }
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      mn[k] = fminf(fminf(p[k], p[4 + k]), p[8 + k]);
      mx[k] = fmaxf(fmaxf(p[k], p[4 + k]), p[8 + k]);
    }
  }
}

// Rays (bvh.rs:51-328): slab pruning, the kind's ray narrow phase.  INSIDE:
// every accepted hit lies inside its own primitive's AABB slab interval (so a
// near-first walk, and pruning inside a synthetic chain, keep the reference's
// answer).
template <int D, int KIND, bool GUARD = false> struct RayPolicy {
  static constexpr bool INSIDE = KIND == KIND_AABB || GUARD;
  RayD<D> r;
  SQ_HD void setup(const float* src) { ray_setup<D>(r, src); }
  SQ_HD bool root(const float* root, float limit) const {
    float te, tx; int ax;
    return slab_exact<D>(r, root, root + D, limit, te, tx, ax);
  }
  SQ_HD void pair(const float* nd, float limit, bool& ha, bool& hb, float& ta, float& tb) const {
    slab_pair<D>(r, nd, limit, ha, hb, ta, tb);
  }
  SQ_HD bool narrow(const float* p, float limit, float& t) const { return sqd::narrow<D, KIND, float, GUARD>(r, p, limit, t); }
};

// Shape casts (bvh.rs:331-414): the slab of the cast ray against each box
// grown by the probe box, [min - probe.max, max - probe.min]
// (shapecast.rs:70-73).  Narrow: AABB leaves answer with the same slab against
// the grown leaf box (its entry); SPHERE leaves with the ball of radius
// r + probe.max[0] (shape_overlap.rs:104-118), with GUARD kept only inside
// the grown box's slab interval; TRIANGLE leaves keep the reference default
// (no contact).
template <int D, int KIND, class R = float, bool GUARD = false> struct CastPolicy {
  static constexpr bool INSIDE = KIND != KIND_SPHERE || GUARD;
  RayD<D, R> r;
  R pmin[D], pmax[D];
  SQ_HD void setup(const R* src) {
    ray_setup<D>(r, src);
#pragma unroll
    for (int k = 0; k < D; ++k) { pmin[k] = src[2 * D + k]; pmax[k] = src[3 * D + k]; }
  }
  SQ_HD bool grown_slab(const R* mn, const R* mx, R limit, R& te, R& tx) const {
    R gmn[D], gmx[D];
#pragma unroll
    for (int k = 0; k < D; ++k) { gmn[k] = mn[k] - pmax[k]; gmx[k] = mx[k] - pmin[k]; }
    int ax;
    return slab_exact<D>(r, gmn, gmx, limit, te, tx, ax);
  }
  SQ_HD bool root(const R* root, R limit) const {
    R te, tx;
    return grown_slab(root, root + D, limit, te, tx);
  }
  SQ_HD void pair(const float* nd, float limit, bool& ha, bool& hb, float& ta, float& tb) const {
    float g[4 * D + 4];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      g[4 * k + 0] = nd[4 * k + 0] - pmax[k]; g[4 * k + 1] = nd[4 * k + 1] - pmin[k];
      g[4 * k + 2] = nd[4 * k + 2] - pmax[k]; g[4 * k + 3] = nd[4 * k + 3] - pmin[k];
    }
    slab_pair<D>(r, g, limit, ha, hb, ta, tb);
  }
  SQ_HD bool narrow(const R* p, R limit, R& t) const {
    if constexpr (KIND == KIND_TRI) {
      return false;
    } else {
      R mn[D], mx[D];
      prim_box<D, KIND>(p, mn, mx);
      R te, tx;
      if constexpr (KIND == KIND_AABB) {
        if (!grown_slab(mn, mx, limit, te, tx)) return false;
        t = te;
        return true;
      } else {
        R oc[D];
#pragma unroll
        for (int k = 0; k < D; ++k) oc[k] = r.o[k] - p[k];
        const R rad = p[D] + pmax[0];
        R b = dotD<D>(oc, r.d);
        R cc = dotD<D>(oc, oc) - rad * rad;
        R disc = b * b - cc;
        if (disc < R(0)) return false;
        R sq = sq_sqrt(disc);
        R tt = -b - sq;
        if (tt < R(0)) tt = -b + sq;
        if (tt < R(0) || tt > limit) return false;
        if constexpr (GUARD) {
          if (!grown_slab(mn, mx, limit, te, tx)) return false;
          if (!(te <= tt && tt <= tx)) return false;
        }
        t = tt;
        return true;
      }
    }
  }
};

// ---------------------------------------------------------------------------
// The walks
// ---------------------------------------------------------------------------

struct SceneView {
  const float* nodes;       // pair nodes, (D+1)*4 floats each
  const float* prims;       // leaf-ordered records
  const uint32_t* ids;      // leaf-ordered original leaf ids
  const float* root;        // [2D] bounds, then the root ref's bits
  const float* quads = nullptr;   // 4-wide records of the pair nodes (walks with QUAD), see quad_records
};

// 4-wide records for the QUAD walk.  Pair node p gets one record of 2 (D+1)
// 16-byte chunks whose first 2D+1 are read: the D box chunks of slot 0, one
// chunk of the four refs {aa, ab, ba, bb}, then the D box chunks of slot 1
// (floats 4 (D+1) ..; the last chunk is padding, so records stay 128-byte
// aligned at D = 3).  Slot s describes child s of p: the two boxes of the
// child's own pair node when it is interior (p's grandchildren), or the
// child's box twice with its leaf ref in the first place and QUAD_EMPTY in the
// second when it is a leaf range.  Every box is a box of the pair layout, an
// exact union of its primitives' boxes, so a walk over them keeps exactly the
// hits of the pair walk (the order-independence argument above) while it
// fetches half as many nodes one after another, in 2D+1 loads per record
// instead of 2 (D+1).  n_nodes bounds the refs read.
constexpr uint32_t QUAD_EMPTY = 0xFFFFFFFFu;   // no second box in this slot (never a ref: at most 2^27 - 1 primitives)

template <int D>
SQ_HD void quad_records(const float* nodes, uint32_t n_nodes, uint32_t p, float* out) {
  constexpr int NF = (D + 1) * 4;
  const float* nd = nodes + (size_t)p * NF;
  uint32_t refs[4];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    float* rec = out + s * NF;
    const uint32_t r = fbits(nd[4 * D + s]);
    if (!(r & LEAF_BIT) && r < n_nodes) {
      const float* cn = nodes + (size_t)r * NF;
      for (int k = 0; k < 4 * D; ++k) rec[k] = cn[k];
      refs[2 * s] = fbits(cn[4 * D + 0]);
      refs[2 * s + 1] = fbits(cn[4 * D + 1]);
    } else {
      for (int k = 0; k < D; ++k) {
        rec[4 * k + 0] = nd[4 * k + 2 * s]; rec[4 * k + 1] = nd[4 * k + 2 * s + 1];
        rec[4 * k + 2] = nd[4 * k + 2 * s]; rec[4 * k + 3] = nd[4 * k + 2 * s + 1];
      }
      refs[2 * s] = r;
      refs[2 * s + 1] = QUAD_EMPTY;
    }
  }
  // slot 0's box chunks are followed by the refs, slot 1's begin at NF
  for (int k = 0; k < 4; ++k) out[4 * D + k] = bitsf(refs[k]);
  for (int k = 0; k < 4; ++k) out[NF + 4 * D + k] = 0.0f;
}

// Stack accessor: entry k of this lane lives at ref[k * stride] / t[k * stride].
struct Stack {
  uint32_t* ref;
  float* t;
  uint32_t stride;
};

struct NearestResult {
  float t;
  uint32_t leaf;   // original id, NO_LEAF on a miss
  uint32_t pos;    // leaf-ordered position of the hit primitive
};

// The two children of a pair node in walk order: whether each is to be
// visited, which comes first, and the entries.  A synthetic pair (one part of
// a reference leaf) is passed through in leaf order, entry -inf, unless every
// accepted hit lies inside its primitive's box (Q::INSIDE), where its boxes
// prune like any other.  REF: child b first (the reference pushes a, then b);
// else the nearer entry first.
template <int D, bool REF, class Q>
SQ_HD void order_pair(const Q& q, const float* nd, float limit, bool& ha, bool& hb, float& ta, float& tb, bool& a_first) {
  if (!Q::INSIDE && synthetic_pair<D>(nd)) {
    ha = hb = true; ta = tb = -SQ_INF; a_first = true;
    return;
  }
  q.pair(nd, limit, ha, hb, ta, tb);
  a_first = REF ? false : ta <= tb;
}

// Nearest (ANY = false) or first accepted hit (ANY = true) of one query, in
// the reference's order (REF) or near-first.
template <int D, int PS, bool ANY, bool FILT, bool REF = false, class Q>
SQ_HD NearestResult walk_nearest(const SceneView& s, const Q& q, float max_toi, const Filter& f, Stack st) {
  NearestResult res{SQ_INF, NO_LEAF, NO_LEAF};
  float limit = max_toi;
  if (!q.root(s.root, limit)) return res;
  uint32_t cur = fbits(s.root[2 * D]);
  uint32_t sp = 0;
  constexpr int NF = (D + 1) * 4;
  while (true) {
    if (!(cur & LEAF_BIT)) {
      const float* nd = s.nodes + (size_t)cur * NF;
      bool ha, hb, a_first; float ta, tb;
      order_pair<D, REF>(q, nd, limit, ha, hb, ta, tb, a_first);
      const uint32_t ra = fbits(nd[4 * D + 0]), rb = fbits(nd[4 * D + 1]);
      if (ha && hb) {
        st.ref[sp * st.stride] = a_first ? rb : ra;
        st.t[sp * st.stride] = a_first ? tb : ta;
        ++sp;
        cur = a_first ? ra : rb;
        continue;
      }
      if (ha) { cur = ra; continue; }
      if (hb) { cur = rb; continue; }
    } else {
      const uint32_t first = cur & FIRST_MASK;
      const uint32_t cnt = ((cur >> 27) & 15u) + 1u;
      for (uint32_t j = 0; j < cnt; ++j) {
        const uint32_t pos = first + j;
        float t;
        if (!q.narrow(s.prims + (size_t)pos * PS, limit, t)) continue;
        const uint32_t leaf = s.ids[pos];
        if (FILT && !accepts(f, leaf)) continue;
        if (ANY) { res.t = t; res.leaf = leaf; res.pos = pos; return res; }
        if (res.leaf == NO_LEAF || hit_before(t, leaf, res.t, res.leaf)) {
          res.t = t; res.leaf = leaf; res.pos = pos; limit = t;
        }
      }
    }
    // pop, skipping entries whose entry lies beyond the current best
    bool got = false;
    while (sp > 0) {
      --sp;
      if (st.t[sp * st.stride] <= limit) { cur = st.ref[sp * st.stride]; got = true; break; }
    }
    if (!got) break;
  }
  return res;
}

// The reference's QueryStats (stats.rs:39-50) of one ray, counted where
// bvh.rs:89-121 (nearest) and bvh.rs:231-257 (all hits) count them: every
// popped node is a node visit and an AABB test, every filter-accepted
// primitive of a passing leaf a narrow test (accepts before the test, as
// there), every narrow test that returns a hit a hit, kept or not.  The
// reference pushes both children of every interior node whose test passes and
// pops every pushed node, so node visits = 1 (the root) + 2 per interior node
// passing; the pieces of an over-full leaf (a synthetic chain) are one
// reference leaf, walked untested in leaf order.  The walk itself is the
// reference's: child b's subtree first, child a pushed with its entry and
// skipped at its pop once beyond the best hit (the pop-time test, see
// order_pair), the provider's own hits.  ALL: no pruning (raycast_all), the
// result is then meaningless.  cnt = {nodes_visited, aabb_tests,
// narrow_tests, hits}.
template <int D, int PS, bool ALL, bool FILT, class Q>
SQ_HD NearestResult walk_counted(const SceneView& s, const Q& q, float max_toi, const Filter& f, Stack st, uint32_t* cnt) {
  NearestResult res{SQ_INF, NO_LEAF, NO_LEAF};
  float limit = max_toi;
  cnt[0] = 1; cnt[1] = 1; cnt[2] = 0; cnt[3] = 0;
  if (!q.root(s.root, limit)) return res;
  uint32_t cur = fbits(s.root[2 * D]);
  uint32_t sp = 0;
  constexpr int NF = (D + 1) * 4;
  while (true) {
    if (!(cur & LEAF_BIT)) {
      const float* nd = s.nodes + (size_t)cur * NF;
      const uint32_t ra = fbits(nd[4 * D + 0]), rb = fbits(nd[4 * D + 1]);
      uint32_t first = ra, second = rb;
      bool hf = true, hs = true;
      float ts = -SQ_INF;
      if (!synthetic_pair<D>(nd)) {
        bool ha, hb;
        float ta, tb;
        q.pair(nd, limit, ha, hb, ta, tb);
        cnt[0] += 2; cnt[1] += 2;
        first = rb; hf = hb;
        second = ra; hs = ha; ts = ta;
      }
      if (hf && hs) {
        st.ref[sp * st.stride] = second;
        st.t[sp * st.stride] = ts;
        ++sp;
        cur = first;
        continue;
      }
      if (hf) { cur = first; continue; }
      if (hs) { cur = second; continue; }
    } else {
      const uint32_t first = cur & FIRST_MASK;
      const uint32_t n = ((cur >> 27) & 15u) + 1u;
      for (uint32_t j = 0; j < n; ++j) {
        const uint32_t pos = first + j;
        const uint32_t leaf = s.ids[pos];
        if (FILT && !accepts(f, leaf)) continue;
        ++cnt[2];
        float t;
        if (!q.narrow(s.prims + (size_t)pos * PS, limit, t)) continue;
        ++cnt[3];
        if (!ALL && (res.leaf == NO_LEAF || hit_before(t, leaf, res.t, res.leaf))) {
          res.t = t; res.leaf = leaf; res.pos = pos; limit = t;
        }
      }
    }
    bool got = false;
    while (sp > 0) {
      --sp;
      if (st.t[sp * st.stride] <= limit) { cur = st.ref[sp * st.stride]; got = true; break; }
    }
    if (!got) break;
  }
  return res;
}

// Visit every leaf range whose node chain passes `pair_pred` (fixed
// predicate, no pruning by results): all-hits, crossings, overlap, shape-cast
// all.  leaf_fn(first, count) handles one leaf range.
template <int D, class PairPred, class LeafFn>
SQ_HD void walk_visit(const SceneView& s, uint32_t root_ref, PairPred&& pair_pred, LeafFn&& leaf_fn, Stack st) {
  uint32_t cur = root_ref;
  uint32_t sp = 0;
  constexpr int NF = (D + 1) * 4;
  while (true) {
    if (!(cur & LEAF_BIT)) {
      const float* nd = s.nodes + (size_t)cur * NF;
      bool ha, hb;
      pair_pred(nd, ha, hb);
      const uint32_t ra = fbits(nd[4 * D + 0]), rb = fbits(nd[4 * D + 1]);
      if (ha && hb) { st.ref[sp * st.stride] = rb; ++sp; cur = ra; continue; }
      if (ha) { cur = ra; continue; }
      if (hb) { cur = rb; continue; }
    } else {
      leaf_fn(cur & FIRST_MASK, ((cur >> 27) & 15u) + 1u);
    }
    if (sp == 0) break;
    --sp;
    cur = st.ref[sp * st.stride];
  }
}

// Box overlap (aabb.rs:101-108): touching counts.
template <int D, class R = float> SQ_HD bool box_overlap(const R* amn, const R* amx, const R* bmn, const R* bmx) {
  bool ok = true;
#pragma unroll
  for (int k = 0; k < D; ++k) ok = ok && !(amn[k] > bmx[k] || amx[k] < bmn[k]);
  return ok;
}

// Exact overlap narrow phase of the built-in kinds: AABB and TRIANGLE by
// their AABB; SPHERE by the nearest point of the region within the radius.
template <int D, int KIND, class R = float> SQ_HD bool overlap_narrow(const R* p, const R* rmn, const R* rmx) {
  if constexpr (KIND == KIND_SPHERE) {
    R d2 = R(0);
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const R c = p[k];
      R v;
      if (c < rmn[k]) v = rmn[k] - c;
      else if (c > rmx[k]) v = c - rmx[k];
      else v = R(0);
      d2 = d2 + v * v;
    }
    return sq_sqrt(d2) <= p[D];
  } else {
    R mn[D], mx[D];
    prim_box<D, KIND>(p, mn, mx);
    return box_overlap<D>(mn, mx, rmn, rmx);
  }
}

// Order-independent minimal work at a fixed limit: every node box that must
// be tested (root + both children of every passing pair) and every primitive
// of every passing leaf.  Counted in the reference layout: a reference leaf
// split into a synthetic chain counts as the one node it is there.
template <int D>
SQ_HD void walk_required(const SceneView& s, const RayD<D>& r, float t_limit, Stack st, uint32_t& n_nodes, uint32_t& n_prims) {
  n_nodes = 1; n_prims = 0;
  float te, tx; int ax;
  if (!slab_exact<D>(r, s.root, s.root + D, t_limit, te, tx, ax)) return;
  walk_visit<D>(s, fbits(s.root[2 * D]),
    [&](const float* nd, bool& ha, bool& hb) {
      const bool synth = fbits(nd[4 * D + 3]) != 0;
      if (synth) { ha = hb = true; return; }
      n_nodes += 2;
      float mn[D], mx[D], e, x; int a;
#pragma unroll
      for (int k = 0; k < D; ++k) { mn[k] = nd[4 * k + 0]; mx[k] = nd[4 * k + 1]; }
      ha = slab_exact<D>(r, mn, mx, t_limit, e, x, a);
#pragma unroll
      for (int k = 0; k < D; ++k) { mn[k] = nd[4 * k + 2]; mx[k] = nd[4 * k + 3]; }
      hb = slab_exact<D>(r, mn, mx, t_limit, e, x, a);
    },
    [&](uint32_t, uint32_t cnt) { n_prims += cnt; }, st);
}

// Sort keys of the CSR queries: (toi in total order, leaf) for hits.
SQ_HD uint64_t hit_key(float t, uint32_t leaf) {
  return ((uint64_t)((uint32_t)total_key(t) ^ 0x80000000u) << 32) | leaf;
}
SQ_HD float key_toi(uint64_t k) {
  int32_t i = (int32_t)((uint32_t)(k >> 32) ^ 0x80000000u);
  i ^= (int32_t)(((uint32_t)(i >> 31)) >> 1);
  return bitsf((uint32_t)i);
}

// ---------------------------------------------------------------------------
// Resumable nearest walk with a ring stack (the pooled kernel's per-lane state)
// ---------------------------------------------------------------------------

struct SpillEntry {
  uint32_t ref;
  uint32_t t;    // bits of the entry distance
};

// The newest S entries of a lane's stack live in LDS slots k & (S-1); when a
// push would overwrite an entry, that older entry moves to the lane's spill
// column (entry k - S at spill[(k - S) * spill_stride]); the matching pop
// brings it back.  S is a power of two.
struct RingStack {
  uint32_t* ref;
  float* t;
  uint32_t stride;          // LDS slot j at ref[j * stride]
  uint32_t S;
  SpillEntry* spill;        // this lane's column
  uint32_t spill_stride;
  SQ_HD void push(uint32_t& sp, uint32_t r, float tt) const {
    const uint32_t slot = (sp & (S - 1)) * stride;
    if (sp >= S) spill[(size_t)(sp - S) * spill_stride] = SpillEntry{ref[slot], fbits(t[slot])};
    ref[slot] = r;
    t[slot] = tt;
    ++sp;
  }
  SQ_HD void pop(uint32_t& sp, uint32_t& r, float& tt) const {
    --sp;
    const uint32_t slot = (sp & (S - 1)) * stride;
    r = ref[slot];
    tt = t[slot];
    if (sp >= S) {
      const SpillEntry e = spill[(size_t)(sp - S) * spill_stride];
      ref[slot] = e.ref;
      t[slot] = bitsf(e.t);
    }
  }
};

struct WalkState {
  uint32_t cur, sp;
  float limit;
  NearestResult res;
};

// Resumable form of walk_visit (the same set of leaf ranges): one call
// descends to a leaf, or stops early when pause() says so, hands the leaf
// range to leaf_fn and pops the next entry.  True when the query has no
// entries left.  Start with cur = root ref, sp = 0.
struct VisitState {
  uint32_t cur, sp;
};

template <int D, class PairPred, class LeafFn, class P>
SQ_HD bool visit_step(const SceneView& s, PairPred&& pair_pred, LeafFn&& leaf_fn, const RingStack& st, VisitState& v,
                      const P& pause) {
  constexpr int NF = (D + 1) * 4;
  float t;
  while (!(v.cur & LEAF_BIT)) {
    const float* nd = s.nodes + (size_t)v.cur * NF;
    bool ha, hb;
    pair_pred(nd, ha, hb);
    const uint32_t ra = fbits(nd[4 * D + 0]), rb = fbits(nd[4 * D + 1]);
    if (ha && hb) {
      st.push(v.sp, rb, 0.0f);
      v.cur = ra;
    } else if (ha) {
      v.cur = ra;
    } else if (hb) {
      v.cur = rb;
    } else {
      if (v.sp == 0) return true;
      st.pop(v.sp, v.cur, t);
    }
    if (pause()) return false;
  }
  leaf_fn(v.cur & FIRST_MASK, ((v.cur >> 27) & 15u) + 1u);
  if (v.sp == 0) return true;
  st.pop(v.sp, v.cur, t);
  return false;
}

// Start a query: false when it misses the root box (the result is a miss).
template <int D, class Q>
SQ_HD bool walk_begin(const SceneView& s, const Q& q, float max_toi, WalkState& w) {
  w.res = NearestResult{SQ_INF, NO_LEAF, NO_LEAF};
  w.limit = max_toi;
  w.sp = 0;
  w.cur = fbits(s.root[2 * D]);
  return q.root(s.root, max_toi);
}

// Work done by a walk: pair nodes fetched and primitives tested.
struct WorkCount {
  uint32_t nodes = 0, prims = 0;
};

// Never pause (one lane, or no wave-level policy).
struct NoPause {
  SQ_HD bool operator()() const { return false; }
};

// One step: descend through pair nodes to the next leaf (the child order_pair
// puts first next, the other pushed with its entry distance; REF walks the
// reference's order, else near-first; when both children miss,
// pop until an entry is no farther than the current best), then test that
// leaf's primitives and pop to the next entry worth visiting.  Returns true
// when the query is finished (nothing left to pop).  pause() is asked after
// every node: when it says so the step returns early with the walk's state
// intact (the wave stops waiting on a few long descents; they resume in the
// next step), so per query the order of work is the same whatever it says.
// COUNT adds the step's work to *wc.  SKIP treats the child refs skip[0..3] as
// missed: subtrees this query has already walked completely (with a limit at
// least the current one, so they hold nothing it has not seen).  QUAD walks
// the 4-wide records: near-first with Q::INSIDE (its records hold the
// synthetic pairs' boxes like any other), or with REF in the reference's
// order over a tree without synthetic pairs.  The REF form visits the
// grandchildren bb, ba, ab, aa of a pair node (child b's subtree first, its
// own child b first, as the reference pushes a then b at every node) and
// tests each against the limit when it is popped, so it visits the leaves the
// pair walk visits, in the same order, with the same limits: a grandchild's
// slab interval lies inside its parent's (parent and child bounds nest and
// each slab operation rounds monotonically), so a grandchild passing at its
// pop means the skipped test of its parent, made at the same limit (nothing
// is tested between the two pops), passes too.
template <int D, int PS, bool ANY, bool FILT, bool COUNT = false, bool SKIP = false, bool QUAD = false, bool REF = false,
          class Q, class P = NoPause>
SQ_HD bool walk_step(const SceneView& s, const Q& q, const Filter& f, const RingStack& st, WalkState& w, WorkCount* wc = nullptr,
                     const uint32_t* skip = nullptr, const P& pause = P()) {
  constexpr int NF = (D + 1) * 4;
  // (any-hit queries never shrink their limit, so they visit the same leaves
  // in any order; without synthetic pairs the 4-wide records prune by the
  // pair layout's own boxes)
  static_assert(!QUAD || Q::INSIDE || REF || ANY, "the 4-wide walk over hits that may leave their boxes walks the reference's order");
  if constexpr (QUAD) {
    // descend over the 4-wide records: the nearest child next, the other hit
    // children pushed farthest first (so they pop nearest first)
    while (!(w.cur & LEAF_BIT)) {
      float nd[2 * NF];
      fetch16<2 * D + 1>(s.quads + (size_t)w.cur * 2 * NF, nd);
      bool h[4];
      float t[4];
      uint32_t r[4];
      q.pair(nd, w.limit, h[0], h[1], t[0], t[1]);
      q.pair(nd + NF, w.limit, h[2], h[3], t[2], t[3]);
#pragma unroll
      for (int k = 0; k < 4; ++k) r[k] = fbits(nd[4 * D + k]);
      h[1] = h[1] && r[1] != QUAD_EMPTY;
      h[3] = h[3] && r[3] != QUAD_EMPTY;
      if constexpr (REF) {
        // slots in the reference's order aa, ab, ba, bb are pushed as found,
        // the last hit one walked next: bb's subtree first, aa's last
        bool have = false;
        uint32_t nr = 0;
        float ntt = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          if (!h[k]) continue;
          if (have) st.push(w.sp, nr, ntt);
          have = true;
          nr = r[k];
          ntt = t[k];
        }
        if (have) {
          w.cur = nr;
        } else {
          bool got = false;
          while (w.sp > 0) {
            uint32_t rr; float tt;
            st.pop(w.sp, rr, tt);
            if (tt <= w.limit) { w.cur = rr; got = true; break; }
          }
          if (!got) return true;
        }
        if (pause()) return false;
        continue;
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) t[k] = h[k] ? t[k] : SQ_INF;
      const uint32_t nh = (uint32_t)h[0] + (uint32_t)h[1] + (uint32_t)h[2] + (uint32_t)h[3];
      // sort the four (t, ref) by t: misses (+inf) go last
      auto cswap = [&](int a, int b) {
        const bool sw = t[b] < t[a];
        const float ta = t[a], tb = t[b];
        const uint32_t ra = r[a], rb = r[b];
        t[a] = sw ? tb : ta; t[b] = sw ? ta : tb;
        r[a] = sw ? rb : ra; r[b] = sw ? ra : rb;
      };
      cswap(0, 1); cswap(2, 3); cswap(0, 2); cswap(1, 3); cswap(1, 2);
      if (nh == 0) {
        bool got = false;
        while (w.sp > 0) {
          uint32_t rr; float tt;
          st.pop(w.sp, rr, tt);
          if (tt <= w.limit) { w.cur = rr; got = true; break; }
        }
        if (!got) return true;
      } else {
        if (nh > 3) st.push(w.sp, r[3], t[3]);
        if (nh > 2) st.push(w.sp, r[2], t[2]);
        if (nh > 1) st.push(w.sp, r[1], t[1]);
        w.cur = r[0];
      }
      if (pause()) return false;
    }
  }
  while (!(w.cur & LEAF_BIT)) {
    float nd[NF];
    fetch16<D + 1>(s.nodes + (size_t)w.cur * NF, nd);
    if constexpr (COUNT) ++wc->nodes;
    bool ha, hb, a_first; float ta, tb;
    order_pair<D, REF>(q, nd, w.limit, ha, hb, ta, tb, a_first);
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
        uint32_t r; float tt;
        st.pop(w.sp, r, tt);
        if (tt <= w.limit) { w.cur = r; got = true; break; }
      }
      if (!got) return true;
    }
    if (pause()) return false;
  }
  const uint32_t first = w.cur & FIRST_MASK;
  const uint32_t cnt = ((w.cur >> 27) & 15u) + 1u;
  if constexpr (COUNT) wc->prims += cnt;
  for (uint32_t j = 0; j < cnt; ++j) {
    const uint32_t pos = first + j;
    float t;
    bool hit;
    if constexpr (PS == 12) {
      float pr[12];
      fetch16<3>(s.prims + (size_t)pos * 12, pr);
      hit = q.narrow(pr, w.limit, t);
    } else {
      hit = q.narrow(s.prims + (size_t)pos * PS, w.limit, t);
    }
    if (!hit) continue;
    const uint32_t leaf = s.ids[pos];
    if (FILT && !accepts(f, leaf)) continue;
    if (ANY) { w.res.t = t; w.res.leaf = leaf; w.res.pos = pos; return true; }
    if (w.res.leaf == NO_LEAF || hit_before(t, leaf, w.res.t, w.res.leaf)) {
      w.res.t = t; w.res.leaf = leaf; w.res.pos = pos; w.limit = t;
    }
  }
  while (w.sp > 0) {
    uint32_t r; float tt;
    st.pop(w.sp, r, tt);
    if (tt <= w.limit) { w.cur = r; return false; }
  }
  return true;
}

// Look-ahead form of the walk in the reference's order (REF): a lane that
// reaches a leaf keeps it pending and goes on descending with the limit it
// has, towards its next leaf in the same order, and stops there; the pending
// leaf is then tested and the descent position is tested again against the
// new limit (popped past when its entry lies beyond it).  Leaves are tested in
// the reference's order, each with the limit the reference tests it with and
// only when its entry passes that limit: entries only grow down the tree, so
// a node passed on the larger limit holds no leaf whose entry passes a smaller
// one unless the node's own entry does, and every entry pushed is tested
// again when popped.  So the walk tests exactly the primitives walk_step
// tests, while lanes that reach a leaf early take their next descent instead
// of waiting for the rest of the wave.  QUAD walks the 4-wide records (trees
// without synthetic pairs), else the pair nodes (synthetic pairs passed
// through in leaf order, as order_pair does).
constexpr uint32_t NO_CUR = 0xFFFFFFFFu;   // no descent position (never a leaf ref: at most 2^27 - 1 primitives)

struct SpecState {
  uint32_t cur, sp, pend;   // descent position (NO_CUR: none left), stack depth, pending leaf (0: none)
  float cur_t, limit;       // the position's entry; the current limit
  NearestResult res;
};

template <int D, class Q>
SQ_HD bool spec_begin(const SceneView& s, const Q& q, float max_toi, SpecState& w) {
  w.res = NearestResult{SQ_INF, NO_LEAF, NO_LEAF};
  w.limit = max_toi;
  w.sp = 0;
  w.pend = 0;
  w.cur = fbits(s.root[2 * D]);
  w.cur_t = -SQ_INF;   // the root is the first position, descended before any leaf test
  return q.root(s.root, max_toi);
}

// Next stack entry whose entry passes `limit` (NO_CUR when none is left).
SQ_HD void pop_next(const RingStack& st, uint32_t& sp, float limit, uint32_t& cur, float& cur_t) {
  cur = NO_CUR;
  while (sp > 0) {
    uint32_t r; float tt;
    st.pop(sp, r, tt);
    if (tt <= limit) { cur = r; cur_t = tt; return; }
  }
}

template <int D, int PS, bool FILT, bool QUAD, class Q, class P = NoPause, bool COUNT = false>
SQ_HD bool walk_step_spec(const SceneView& s, const Q& q, const Filter& f, const RingStack& st, SpecState& w,
                          const P& pause = P(), WorkCount* wc = nullptr) {
  constexpr int NF = (D + 1) * 4;
  if (w.pend == 0 && w.cur != NO_CUR && (w.cur & LEAF_BIT)) {
    w.pend = w.cur;
    pop_next(st, w.sp, w.limit, w.cur, w.cur_t);
  }
  // descend: to the next leaf, and past a pending one to the leaf after it
  while (w.cur != NO_CUR && !(w.cur & LEAF_BIT)) {
    if constexpr (QUAD) {
      float nd[2 * NF];
      fetch16<2 * D + 1>(s.quads + (size_t)w.cur * 2 * NF, nd);
      bool h[4];
      float t[4];
      uint32_t r[4];
      q.pair(nd, w.limit, h[0], h[1], t[0], t[1]);
      q.pair(nd + NF, w.limit, h[2], h[3], t[2], t[3]);
#pragma unroll
      for (int k = 0; k < 4; ++k) r[k] = fbits(nd[4 * D + k]);
      h[1] = h[1] && r[1] != QUAD_EMPTY;
      h[3] = h[3] && r[3] != QUAD_EMPTY;
      // aa, ab, ba pushed as found, the last hit one walked next (bb first)
      bool have = false;
      uint32_t nr = 0;
      float ntt = 0.0f;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (!h[k]) continue;
        if (have) st.push(w.sp, nr, ntt);
        have = true;
        nr = r[k];
        ntt = t[k];
      }
      if (have) { w.cur = nr; w.cur_t = ntt; }
      else pop_next(st, w.sp, w.limit, w.cur, w.cur_t);
    } else {
      float nd[NF];
      fetch16<D + 1>(s.nodes + (size_t)w.cur * NF, nd);
      bool ha, hb, a_first; float ta, tb;
      order_pair<D, true>(q, nd, w.limit, ha, hb, ta, tb, a_first);
      const uint32_t ra = fbits(nd[4 * D + 0]), rb = fbits(nd[4 * D + 1]);
      if (ha && hb) {
        st.push(w.sp, a_first ? rb : ra, a_first ? tb : ta);
        w.cur = a_first ? ra : rb;
        w.cur_t = a_first ? ta : tb;
      } else if (ha) { w.cur = ra; w.cur_t = ta; }
      else if (hb) { w.cur = rb; w.cur_t = tb; }
      else pop_next(st, w.sp, w.limit, w.cur, w.cur_t);
    }
    if (w.pend == 0 && w.cur != NO_CUR && (w.cur & LEAF_BIT)) {
      w.pend = w.cur;
      pop_next(st, w.sp, w.limit, w.cur, w.cur_t);
    }
    if (pause()) break;
  }
  if (w.pend == 0) return w.cur == NO_CUR;
  const uint32_t first = w.pend & FIRST_MASK;
  const uint32_t cnt = ((w.pend >> 27) & 15u) + 1u;
  w.pend = 0;
  if constexpr (COUNT) wc->prims += cnt;
  if constexpr (PS == 12) {
    // triangles: the next record is fetched while this one is tested; the
    // id rides in the record (v0.xyz, id)
    float pr[12], nx[12];
    fetch16<3>(s.prims + (size_t)first * 12, pr);
    for (uint32_t j = 0; j < cnt; ++j) {
      const uint32_t pos = first + j;
      if (j + 1 < cnt) fetch16<3>(s.prims + (size_t)(pos + 1) * 12, nx);
      float t;
      if (q.narrow(pr, w.limit, t)) {
        const uint32_t leaf = fbits(pr[3]);
        if (!FILT || accepts(f, leaf)) {
          if (w.res.leaf == NO_LEAF || hit_before(t, leaf, w.res.t, w.res.leaf)) {
            w.res.t = t; w.res.leaf = leaf; w.res.pos = pos; w.limit = t;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < 12; ++k) pr[k] = nx[k];
    }
  } else {
    for (uint32_t j = 0; j < cnt; ++j) {
      const uint32_t pos = first + j;
      float t;
      if (!q.narrow(s.prims + (size_t)pos * PS, w.limit, t)) continue;
      const uint32_t leaf = s.ids[pos];
      if (FILT && !accepts(f, leaf)) continue;
      if (w.res.leaf == NO_LEAF || hit_before(t, leaf, w.res.t, w.res.leaf)) {
        w.res.t = t; w.res.leaf = leaf; w.res.pos = pos; w.limit = t;
      }
    }
  }
  if (w.cur != NO_CUR && !(w.cur_t <= w.limit)) pop_next(st, w.sp, w.limit, w.cur, w.cur_t);
  return w.cur == NO_CUR;
}

// ---------------------------------------------------------------------------
// Two-level walk (crates/spatial-query/src/accel/tlas.rs) over built-in
// instance kinds.  Instance records by id, XS floats each:
//   INST_BOX  (any D): R[D*D] row-major, t[D], half[D]          two_level.rs:70-113
//   INST_MESH (D = 3): R[9], t[3], 1/scale, mesh id (bits)        a uniformly
//                      scaled triangle mesh answered by its own device BVH.
//   INST_BALL (D = 3): R[9], t[3], radius                          a ball at the
//                      local origin (BallInstances, shape_overlap.rs:288-345):
//                      the one kind that also answers shape casts and an exact
//                      overlap test (the others keep the trait defaults).
// With GUARD every kind keeps an instance hit only when its toi lies in the
// slab interval of the instance's world AABB, which makes the top-level result
// independent of traversal order (the near-first walk); without it the hit is
// the reference's (tlas.rs:459-481) and nearest queries walk the top level in
// the reference's order.
// ---------------------------------------------------------------------------

enum { INST_BOX = 0, INST_MESH = 1, INST_BALL = 2 };

// Isometry::inverse_transform_point / _vector (isometry.rs:70-78, 113-123):
// rotate_transpose(o - t) and rotate_transpose(d), each accumulating from 0.
template <int D, class S = float>
SQ_HD void iso_inverse(const S* R, const S* t, const RayD<D, S>& r, S* o_l, S* d_l) {
  S v[D];
#pragma unroll
  for (int j = 0; j < D; ++j) v[j] = r.o[j] - t[j];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    S a = S(0), b = S(0);
#pragma unroll
    for (int j = 0; j < D; ++j) { a = a + R[j * D + i] * v[j]; b = b + R[j * D + i] * r.d[j]; }
    o_l[i] = a; d_l[i] = b;
  }
}

// Isometry::transform_vector (rotate, isometry.rs:99-110).
template <int D> SQ_HD void iso_rotate(const float* R, const float* n, float* out) {
#pragma unroll
  for (int i = 0; i < D; ++i) {
    float a = 0.0f;
#pragma unroll
    for (int j = 0; j < D; ++j) a = a + R[i * D + j] * n[j];
    out[i] = a;
  }
}

// ray_local_box (two_level.rs:81-113): a local box [-half, half]; parallel
// threshold 1e-9; a hit needs an entry axis, t_min > 1e-6 and t_min <= max_toi.
// (The literals are the Scalar's: 1e-9 and 1e-6 rounded to float, or double.)
template <int D, class S = float>
SQ_HD bool local_box(const S* o, const S* d, const S* half, S max_toi, S& t, int& axis_out) {
  const S par = S(sizeof(S) == 4 ? (double)1.0e-9f : 1.0e-9);
  const S tiny = S(sizeof(S) == 4 ? (double)1.0e-6f : 1.0e-6);
  S t_min = S(0), t_max = max_toi;
  int axis = -1;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    if (sq_abs(d[k]) < par) {
      if (o[k] < -half[k] || o[k] > half[k]) return false;
    } else {
      const S inv = SQ_DIV_RN(S(1), d[k]);
      S t1 = (-half[k] - o[k]) * inv;
      S t2 = (half[k] - o[k]) * inv;
      if (t1 > t2) { S tmp = t1; t1 = t2; t2 = tmp; }
      if (t1 > t_min) { t_min = t1; axis = k; }
      if (t2 < t_max) t_max = t2;
      if (t_min > t_max) return false;
    }
  }
  if (!(axis >= 0 && t_min > tiny && t_min <= max_toi)) return false;
  t = t_min; axis_out = axis;
  return true;
}

// ray_ball (shape_overlap.rs:120-138) about centre c: the entry root, else the
// exit root; rejects t < 0 or t > max_toi.  With c = NULL the centre is the
// origin (a local ray against a ball instance).  Scalar-generic (the f64
// feature instantiates it on double).
template <int D, class S = float> SQ_HD bool ray_ball(const S* o, const S* d, const S* c, S rad, S max_toi, S& t) {
  S oc[D];
#pragma unroll
  for (int k = 0; k < D; ++k) oc[k] = c ? o[k] - c[k] : o[k];
  const S b = dotD<D>(oc, d);
  const S cc = dotD<D>(oc, oc) - rad * rad;
  const S disc = b * b - cc;
  if (disc < S(0)) return false;
  const S sq = sq_sqrt(disc);
  S tt = -b - sq;
  if (tt < S(0)) tt = -b + sq;
  if (tt < S(0) || tt > max_toi) return false;
  t = tt;
  return true;
}

struct TlasView {
  SceneView top;            // pair nodes over instance world boxes; top.ids are instance ids
  const float* xf;          // instance records by id
  const float* world;       // instance world boxes by id, [n][2D]
  const SceneView* blas;    // per mesh (INST_MESH)
  uint32_t xs;              // floats per instance record
};

struct TlasHit {
  float t;
  uint32_t inst;   // NO_LEAF on a miss
  uint32_t pos;    // BLAS position of the triangle (mesh) or the entry axis (box)
  uint32_t sub;    // triangle id within the mesh, NO_LEAF for boxes
};

// test_instance (tlas.rs:459-481); GUARD adds the instance guard.  A mesh
// instance answers with its mesh's nearest hit: the reference's walk of the
// mesh BVH (unguarded triangles, the reference's order), or with GUARD the
// guarded near-first walk.
template <int D, int IK, bool GUARD = false>
SQ_HD bool instance_test(const TlasView& tv, const RayD<D>& r, uint32_t id, float limit, Stack st_bot, TlasHit& h) {
  const float* x = tv.xf + (size_t)id * tv.xs;
  float ol[D], dl[D];
  iso_inverse<D>(x, x + D * D, r, ol, dl);
  float t;
  if constexpr (IK == INST_BOX) {
    int ax;
    if (!local_box<D>(ol, dl, x + D * D + D, limit, t, ax)) return false;
    h.pos = (uint32_t)ax; h.sub = NO_LEAF;
  } else if constexpr (IK == INST_BALL) {
    if (!ray_ball<D>(ol, dl, (const float*)nullptr, x[D * D + D], limit, t)) return false;
    h.pos = 0; h.sub = NO_LEAF;
  } else {
    const float inv_s = x[D * D + D];
    const uint32_t mesh = fbits(x[D * D + D + 1]);
    float src[2 * D];
#pragma unroll
    for (int k = 0; k < D; ++k) { src[k] = ol[k] * inv_s; src[D + k] = dl[k] * inv_s; }
    RayPolicy<D, KIND_TRI, GUARD> pol;
    pol.setup(src);
    Filter nof{nullptr, ~0ull, -1};
    NearestResult b = walk_nearest<D, Prim<D, KIND_TRI>::STRIDE, false, false, !GUARD>(tv.blas[mesh], pol, limit, nof, st_bot);
    if (b.leaf == NO_LEAF) return false;
    t = b.t; h.pos = b.pos; h.sub = b.leaf;
  }
  if constexpr (GUARD) {
    const float* w = tv.world + (size_t)id * 2 * D;
    float te, tx; int a;
    if (!slab_exact<D>(r, w, w + D, limit, te, tx, a)) return false;
    if (!(te <= t && t <= tx)) return false;
  }
  h.t = t; h.inst = id;
  return true;
}

// The two top-level children of a pair node in walk order (as order_pair):
// EXACT (unguarded instance tests) passes synthetic pairs through in leaf
// order and walks nearest queries in the reference's order (b first); else
// near-first over the guarded tests.
template <int D, bool EXACT, bool ANY>
SQ_HD void order_top(const RayD<D>& r, const float* nd, float limit, bool& ha, bool& hb, float& ta, float& tb, bool& a_first) {
  if (EXACT && synthetic_pair<D>(nd)) {
    ha = hb = true; ta = tb = -SQ_INF; a_first = true;
    return;
  }
  slab_pair<D>(r, nd, limit, ha, hb, ta, tb);
  a_first = (EXACT && !ANY) ? false : ta <= tb;
}

// Tlas::raycast_nearest (tlas.rs:104-154); ANY stops at the first hit.
// EXACT: unguarded instance tests in the reference's order; else guarded,
// near-first.
template <int D, int IK, bool ANY, bool EXACT = true>
SQ_HD TlasHit walk_tlas(const TlasView& tv, const RayD<D>& r, float max_toi, Stack st, Stack st_bot) {
  TlasHit res{SQ_INF, NO_LEAF, NO_LEAF, NO_LEAF};
  float limit = max_toi;
  const SceneView& s = tv.top;
  {
    float te, tx; int ax;
    if (!slab_exact<D>(r, s.root, s.root + D, limit, te, tx, ax)) return res;
  }
  uint32_t cur = fbits(s.root[2 * D]);
  uint32_t sp = 0;
  constexpr int NF = (D + 1) * 4;
  while (true) {
    if (!(cur & LEAF_BIT)) {
      const float* nd = s.nodes + (size_t)cur * NF;
      bool ha, hb, a_first; float ta, tb;
      order_top<D, EXACT, ANY>(r, nd, limit, ha, hb, ta, tb, a_first);
      const uint32_t ra = fbits(nd[4 * D + 0]), rb = fbits(nd[4 * D + 1]);
      if (ha && hb) {
        st.ref[sp * st.stride] = a_first ? rb : ra;
        st.t[sp * st.stride] = a_first ? tb : ta;
        ++sp;
        cur = a_first ? ra : rb;
        continue;
      }
      if (ha) { cur = ra; continue; }
      if (hb) { cur = rb; continue; }
    } else {
      const uint32_t first = cur & FIRST_MASK;
      const uint32_t cnt = ((cur >> 27) & 15u) + 1u;
      for (uint32_t j = 0; j < cnt; ++j) {
        TlasHit h;
        if (!instance_test<D, IK, !EXACT>(tv, r, s.ids[first + j], limit, st_bot, h)) continue;
        if (ANY) return h;
        if (res.inst == NO_LEAF || hit_before(h.t, h.inst, res.t, res.inst)) { res = h; limit = h.t; }
      }
    }
    bool got = false;
    while (sp > 0) {
      --sp;
      if (st.t[sp * st.stride] <= limit) { cur = st.ref[sp * st.stride]; got = true; break; }
    }
    if (!got) break;
  }
  return res;
}

// World normal of a final TLAS hit: the local normal lifted by R (tlas.rs:478).
template <int D, int IK>
SQ_HD void tlas_normal(const TlasView& tv, const RayD<D>& r, const TlasHit& h, float limit, float* n) {
  const float* x = tv.xf + (size_t)h.inst * tv.xs;
  float ol[D], dl[D], nl[D];
  iso_inverse<D>(x, x + D * D, r, ol, dl);
#pragma unroll
  for (int k = 0; k < D; ++k) nl[k] = 0.0f;
  if constexpr (IK == INST_BOX) {
#pragma unroll
    for (int k = 0; k < D; ++k) nl[k] = ((uint32_t)k == h.pos) ? (dl[k] > 0.0f ? -1.0f : 1.0f) : 0.0f;
  } else if constexpr (IK == INST_BALL) {
    // normalize_or_zero(local ray at t) (shape_overlap.rs:136)
    float v[D];
#pragma unroll
    for (int k = 0; k < D; ++k) v[k] = ol[k] + dl[k] * h.t;
    const float len = SQ_SQRT_RN(dotD<D>(v, v));
    if (len > F32_EPSILON) {
      const float s = SQ_DIV_RN(1.0f, len);
#pragma unroll
      for (int k = 0; k < D; ++k) nl[k] = v[k] * s;
    }
  } else {
    const float inv_s = x[D * D + D];
    const uint32_t mesh = fbits(x[D * D + D + 1]);
    float src[2 * D];
#pragma unroll
    for (int k = 0; k < D; ++k) { src[k] = ol[k] * inv_s; src[D + k] = dl[k] * inv_s; }
    RayD<D> mr;
    ray_setup<D>(mr, src);
    hit_normal<D, KIND_TRI>(mr, tv.blas[mesh].prims + (size_t)h.pos * Prim<D, KIND_TRI>::STRIDE, h.t, limit, nl);
  }
  iso_rotate<D>(x, nl, n);
}

// Tlas shape casts (tlas.rs:306-391): the swept probe (radius = the cast
// box's max x, as the built-in SPHERE cast) against ball instances, the ray
// against the radius sum about the instance's centre (shape_overlap.rs:
// 326-337), with GUARD kept only inside the probe's slab interval against the
// instance's world AABB.  The other kinds answer no cast
// (InstancedGeometry::test_shape_cast's default, instance.rs:72-87).
template <int D, int IK, bool GUARD = false>
SQ_HD bool instance_cast(const TlasView& tv, const CastPolicy<D, KIND_AABB>& pol, uint32_t id, float limit, float& t) {
  if constexpr (IK != INST_BALL) {
    return false;
  } else {
    const float* x = tv.xf + (size_t)id * tv.xs;
    if (!ray_ball<D>(pol.r.o, pol.r.d, x + D * D, x[D * D + D] + pol.pmax[0], limit, t)) return false;
    if constexpr (!GUARD) return true;
    const float* w = tv.world + (size_t)id * 2 * D;
    float te, tx;
    if (!pol.grown_slab(w, w + D, limit, te, tx)) return false;
    return te <= t && t <= tx;
  }
}

// InstancedGeometry::test_overlap: exact for ball instances (the distance
// from the centre to the region, shape_overlap.rs:59-72,339-341), the
// default (true) for the other kinds.
template <int D, int IK>
SQ_HD bool instance_overlap(const TlasView& tv, uint32_t id, const float* rmn, const float* rmx) {
  if constexpr (IK != INST_BALL) {
    return true;
  } else {
    const float* x = tv.xf + (size_t)id * tv.xs;
    float p[D + 1];
#pragma unroll
    for (int k = 0; k < D; ++k) p[k] = x[D * D + k];
    p[D] = x[D * D + D];
    return overlap_narrow<D, KIND_SPHERE>(p, rmn, rmx);
  }
}

}  // namespace sqd
