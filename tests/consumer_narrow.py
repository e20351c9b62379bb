"""Consumer narrow phases for the tests.  Each source compiles both as HIP
device code (hipRTC, inside the engine) and, with __device__ defined away, as
C++ for the CPU oracle (g++), so the engine and the oracle run the same
arithmetic; the oracle then walks the reference's LIFO order with it."""
import ctypes
import hashlib
import os
import subprocess

# The built-in SPHERE narrow phase written as a consumer: differential.rs:61-82
# (exit root when the origin is inside), dots as ((0 + a0 b0) + a1 b1) + ...
BALL = r"""
__device__ bool sq_narrow(const float* p, const float* o, const float* d, float max_toi, float* toi) {
  float oc[SQ_RTC_D];
  for (int k = 0; k < SQ_RTC_D; ++k) oc[k] = o[k] - p[k];
  float b = 0.0f, cc = 0.0f;
  for (int k = 0; k < SQ_RTC_D; ++k) b = b + oc[k] * d[k];
  for (int k = 0; k < SQ_RTC_D; ++k) cc = cc + oc[k] * oc[k];
  cc = cc - p[SQ_RTC_D] * p[SQ_RTC_D];
  const float disc = b * b - cc;
  if (disc < 0.0f) return false;
  const float sq = __builtin_sqrtf(disc);
  float t = -b - sq;
  if (t < 0.0f) t = -b + sq;
  if (t < 0.0f || t > max_toi) return false;
  *toi = t;
  return true;
}
"""

# An axis-aligned ellipsoid (centre, semi-axes): the unit sphere after scaling
# by 1/a, nearest root of A t^2 + 2 B t + C = 0.
ELLIPSOID = r"""
__device__ bool sq_narrow(const float* p, const float* o, const float* d, float max_toi, float* toi) {
  float u[3], v[3];
  for (int k = 0; k < 3; ++k) { u[k] = (o[k] - p[k]) / p[3 + k]; v[k] = d[k] / p[3 + k]; }
  const float A = ((0.0f + v[0] * v[0]) + v[1] * v[1]) + v[2] * v[2];
  const float B = ((0.0f + u[0] * v[0]) + u[1] * v[1]) + u[2] * v[2];
  const float C = (((0.0f + u[0] * u[0]) + u[1] * u[1]) + u[2] * u[2]) - 1.0f;
  const float disc = B * B - A * C;
  if (disc < 0.0f) return false;
  const float s = __builtin_sqrtf(disc);
  float t = (-B - s) / A;
  if (t < 0.0f) t = (-B + s) / A;
  if (t < 0.0f || t > max_toi) return false;
  *toi = t;
  return true;
}
"""


# BALL with the built-in SPHERE kind's shape cast (a ball probe of radius the
# probe box's max x: the ray against the radius sum, shape_overlap.rs:96-110)
# and exact overlap (distance from the centre to the region <= r,
# shape_overlap.rs:59-72,121-124), in the same operation order.
BALL_FULL = BALL + r"""
#define SQ_NARROW_CAST 1
__device__ bool sq_narrow_cast(const float* p, const float* o, const float* d, const float* pmin, const float* pmax,
                               float max_toi, float* toi) {
  float oc[SQ_RTC_D];
  for (int k = 0; k < SQ_RTC_D; ++k) oc[k] = o[k] - p[k];
  const float rad = p[SQ_RTC_D] + pmax[0];
  float b = 0.0f, cc = 0.0f;
  for (int k = 0; k < SQ_RTC_D; ++k) b = b + oc[k] * d[k];
  for (int k = 0; k < SQ_RTC_D; ++k) cc = cc + oc[k] * oc[k];
  cc = cc - rad * rad;
  const float disc = b * b - cc;
  if (disc < 0.0f) return false;
  const float sq = __builtin_sqrtf(disc);
  float t = -b - sq;
  if (t < 0.0f) t = -b + sq;
  if (t < 0.0f || t > max_toi) return false;
  *toi = t;
  return true;
}
#define SQ_NARROW_OVERLAP 1
__device__ bool sq_narrow_overlap(const float* p, const float* mn, const float* mx) {
  float d2 = 0.0f;
  for (int k = 0; k < SQ_RTC_D; ++k) {
    float v = 0.0f;
    if (p[k] < mn[k]) v = mn[k] - p[k];
    else if (p[k] > mx[k]) v = p[k] - mx[k];
    d2 = d2 + v * v;
  }
  return __builtin_sqrtf(d2) <= p[SQ_RTC_D];
}
"""

# The spheres of crossings.rs:38-111 (Shells): both roots of the unbounded ray
# are crossings; the nearest hit is the first root at or after the origin.
# Records: centre [D], radius.
SHELL = r"""
#define SQ_NARROW_CROSSINGS 1
__device__ bool shell_roots(const float* p, const float* o, const float* d, float* t0, float* t1) {
  float oc[SQ_RTC_D];
  for (int k = 0; k < SQ_RTC_D; ++k) oc[k] = o[k] - p[k];
  float b = 0.0f, l = 0.0f;
  for (int k = 0; k < SQ_RTC_D; ++k) b = b + oc[k] * d[k];
  for (int k = 0; k < SQ_RTC_D; ++k) l = l + oc[k] * oc[k];
  const float disc = b * b - (l - p[SQ_RTC_D] * p[SQ_RTC_D]);
  if (disc < 0.0f) return false;
  const float sq = __builtin_sqrtf(disc);
  *t0 = -b - sq;
  *t1 = -b + sq;
  return true;
}
__device__ bool sq_narrow(const float* p, const float* o, const float* d, float max_toi, float* toi) {
  float t0, t1;
  if (!shell_roots(p, o, d, &t0, &t1)) return false;
  const float t = t0 >= 0.0f ? t0 : t1;
  if (t < 0.0f || t > max_toi) return false;
  *toi = t;
  return true;
}
template <class Out>
__device__ void sq_narrow_crossings(const float* p, const float* o, const float* d, float max_toi, Out& out) {
  float t0, t1;
  if (!shell_roots(p, o, d, &t0, &t1)) return;
  if (t0 >= 0.0f && t0 <= max_toi) out(t0);
  if (t1 >= 0.0f && t1 <= max_toi) out(t1);
}
"""

# The implicit blobs of field.rs:40-150 (D = 3): the field |x - c|^2 - r0^2 +
# amp (sin(f x0) + sin(f x1) + sin(f x2)), marched from the origin in steps
# of 0.05 to max_toi, every sign change bisected 40 times to a crossing; the
# nearest hit is the first crossing.  Records: centre [3], r0, amp, freq.
# The sine is a polynomial of plain operations (reduction to [-pi, pi], odd
# Taylor terms to x^17), so the device and g++ builds compute the same bits.
BLOB = r"""
#define SQ_NARROW_CROSSINGS 1
__device__ float blob_sin(float x) {
  const float k = __builtin_floorf(x * 0.15915494f + 0.5f);
  const float r = (x - k * 6.28125f) - k * 0.0019353072f;
  const float r2 = r * r;
  float s = 1.0f / 355687428096000.0f;
  s = s * r2 - 1.0f / 1307674368000.0f;
  s = s * r2 + 1.0f / 6227020800.0f;
  s = s * r2 - 1.0f / 39916800.0f;
  s = s * r2 + 1.0f / 362880.0f;
  s = s * r2 - 1.0f / 5040.0f;
  s = s * r2 + 1.0f / 120.0f;
  s = s * r2 - 1.0f / 6.0f;
  s = s * r2 + 1.0f;
  return s * r;
}
__device__ float blob_field(const float* p, const float* o, const float* d, float t) {
  float x[3];
  for (int k = 0; k < 3; ++k) x[k] = (o[k] + d[k] * t) - p[k];
  const float base = (((0.0f + x[0] * x[0]) + x[1] * x[1]) + x[2] * x[2]) - p[3] * p[3];
  const float bump = p[4] * ((blob_sin(p[5] * x[0]) + blob_sin(p[5] * x[1])) + blob_sin(p[5] * x[2]));
  return base + bump;
}
template <class Out>
__device__ void sq_narrow_crossings(const float* p, const float* o, const float* d, float max_toi, Out& out) {
  const float step = 0.05f;
  float t = 0.0f;
  float fp = blob_field(p, o, d, t);
  while (t < max_toi) {
    float tn = t + step;
    if (tn > max_toi) tn = max_toi;
    const float fn = blob_field(p, o, d, tn);
    if (fp != 0.0f && ((fp < 0.0f) != (fn < 0.0f))) {
      float a = t, c = tn, fa = fp;
      for (int i = 0; i < 40; ++i) {
        const float m = 0.5f * (a + c);
        const float fm = blob_field(p, o, d, m);
        if ((fa < 0.0f) != (fm < 0.0f)) c = m;
        else { a = m; fa = fm; }
      }
      out(0.5f * (a + c));
    }
    if (tn >= max_toi) break;
    t = tn;
    fp = fn;
  }
}
struct BlobFirst {
  float t;
  bool hit;
  __device__ void operator()(float x) { if (!hit) { t = x; hit = true; } }
};
__device__ bool sq_narrow(const float* p, const float* o, const float* d, float max_toi, float* toi) {
  BlobFirst f{0.0f, false};
  sq_narrow_crossings(p, o, d, max_toi, f);
  if (!f.hit) return false;
  *toi = f.t;
  return true;
}
"""


def cpu_entry_points(source: str, dim: int, pf: int, workdir: str):
    """The consumer source built for the CPU (g++ -O2 -ffp-contract=off) as a
    shared object; returns (library, address of its narrow entry point,
    address of its crossings entry point or 0 when the source has none)."""
    code = (f"#define __device__\n#define SQ_RTC_D {dim}\n#define SQ_RTC_PF {pf}\n" + source +
            '\nextern "C" int sq_narrow_c(const float* p, const float* o, const float* d, float m, float* t) '
            "{ return sq_narrow(p, o, d, m, t) ? 1 : 0; }\n"
            "#ifdef SQ_NARROW_CAST\n"
            'extern "C" int sq_cast_c(const float* p, const float* o, const float* d, const float* a, const float* b, '
            "float m, float* t) { return sq_narrow_cast(p, o, d, a, b, m, t) ? 1 : 0; }\n"
            "#endif\n"
            "#ifdef SQ_NARROW_OVERLAP\n"
            'extern "C" int sq_overlap_c(const float* p, const float* a, const float* b) '
            "{ return sq_narrow_overlap(p, a, b) ? 1 : 0; }\n"
            "#endif\n"
            "#ifdef SQ_NARROW_CROSSINGS\n"
            "struct SqEmitC { void (*f)(void*, float); void* c; void operator()(float t) const { f(c, t); } };\n"
            'extern "C" void sq_crossings_c(const float* p, const float* o, const float* d, float m, '
            "void (*f)(void*, float), void* c) { SqEmitC e{f, c}; sq_narrow_crossings(p, o, d, m, e); }\n"
            "#endif\n")
    # one file name per source and shape: dlopen returns an already loaded
    # library for a path it has seen
    tag = hashlib.sha1(code.encode()).hexdigest()[:12]
    src = os.path.join(workdir, f"narrow_{tag}.cpp")
    so = os.path.join(workdir, f"narrow_{tag}.so")
    with open(src, "w") as f:
        f.write(code)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-o", so, src], check=True)
    lib = ctypes.CDLL(so)
    x = ctypes.cast(lib.sq_crossings_c, ctypes.c_void_p).value if hasattr(lib, "sq_crossings_c") else 0
    return lib, ctypes.cast(lib.sq_narrow_c, ctypes.c_void_p).value, x


def cast_overlap_entry_points(lib):
    """Addresses of the shape-cast and overlap entry points of a library from
    cpu_entry_points (0 where the source defines none)."""
    def addr(name):
        return ctypes.cast(getattr(lib, name), ctypes.c_void_p).value if hasattr(lib, name) else 0
    return addr("sq_cast_c"), addr("sq_overlap_c")


def compile_cpu(source: str, dim: int, pf: int, workdir: str):
    """As cpu_entry_points: (library, address of its narrow entry point)."""
    lib, addr, _ = cpu_entry_points(source, dim, pf, workdir)
    return lib, addr


# Consumer narrow phases in double (sq_narrow64, the f64 feature).  BALL64 is
# the built-in f64 SPHERE narrow phase in the same operation order
# (differential.rs:61-82); ELLIPSOID64 the axis-aligned ellipsoid above in double.
BALL64 = r"""
__device__ bool sq_narrow64(const double* p, const double* o, const double* d, double max_toi, double* toi) {
  double oc[SQ_RTC_D];
  for (int k = 0; k < SQ_RTC_D; ++k) oc[k] = o[k] - p[k];
  double b = 0.0, cc = 0.0;
  for (int k = 0; k < SQ_RTC_D; ++k) b = b + oc[k] * d[k];
  for (int k = 0; k < SQ_RTC_D; ++k) cc = cc + oc[k] * oc[k];
  cc = cc - p[SQ_RTC_D] * p[SQ_RTC_D];
  const double disc = b * b - cc;
  if (disc < 0.0) return false;
  const double sq = __builtin_sqrt(disc);
  double t = -b - sq;
  if (t < 0.0) t = -b + sq;
  if (t < 0.0 || t > max_toi) return false;
  *toi = t;
  return true;
}
"""

ELLIPSOID64 = r"""
__device__ bool sq_narrow64(const double* p, const double* o, const double* d, double max_toi, double* toi) {
  double u[3], v[3];
  for (int k = 0; k < 3; ++k) { u[k] = (o[k] - p[k]) / p[3 + k]; v[k] = d[k] / p[3 + k]; }
  const double A = ((0.0 + v[0] * v[0]) + v[1] * v[1]) + v[2] * v[2];
  const double B = ((0.0 + u[0] * v[0]) + u[1] * v[1]) + u[2] * v[2];
  const double C = (((0.0 + u[0] * u[0]) + u[1] * u[1]) + u[2] * u[2]) - 1.0;
  const double disc = B * B - A * C;
  if (disc < 0.0) return false;
  const double s = __builtin_sqrt(disc);
  double t = (-B - s) / A;
  if (t < 0.0) t = (-B + s) / A;
  if (t < 0.0 || t > max_toi) return false;
  *toi = t;
  return true;
}
"""


def compile_cpu64(source: str, dim: int, pd: int, workdir: str):
    """An f64 consumer source built for the CPU (g++ -O2 -ffp-contract=off);
    returns (library, address of its int sq_narrow64_c(...) entry point) for
    oracle.Consumer64."""
    code = (f"#define __device__\n#define SQ_RTC_D {dim}\n#define SQ_RTC_PF {pd}\n" + source +
            '\nextern "C" int sq_narrow64_c(const double* p, const double* o, const double* d, double m, double* t) '
            "{ return sq_narrow64(p, o, d, m, t) ? 1 : 0; }\n"
            "#ifdef SQ_NARROW64_CROSSINGS\n"
            "struct SqEmit64 { void (*f)(void*, double); void* c; void operator()(double t) const { f(c, t); } };\n"
            'extern "C" void sq_crossings64_c(const double* p, const double* o, const double* d, double m, '
            "void (*f)(void*, double), void* c) { SqEmit64 e{f, c}; sq_narrow64_crossings(p, o, d, m, e); }\n"
            "#endif\n")
    tag = hashlib.sha1(code.encode()).hexdigest()[:12]
    src = os.path.join(workdir, f"narrow64_{tag}.cpp")
    so = os.path.join(workdir, f"narrow64_{tag}.so")
    with open(src, "w") as f:
        f.write(code)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-o", so, src], check=True)
    lib = ctypes.CDLL(so)
    return lib, ctypes.cast(lib.sq_narrow64_c, ctypes.c_void_p).value


# The shells of crossings.rs:38-111 in double: both roots of the unbounded ray
# are crossings; the nearest hit is the first root at or after the origin.
# Records: centre [D], radius.
SHELL64 = r"""
#define SQ_NARROW64_CROSSINGS 1
__device__ bool shell64_roots(const double* p, const double* o, const double* d, double* t0, double* t1) {
  double oc[SQ_RTC_D];
  for (int k = 0; k < SQ_RTC_D; ++k) oc[k] = o[k] - p[k];
  double b = 0.0, l = 0.0;
  for (int k = 0; k < SQ_RTC_D; ++k) b = b + oc[k] * d[k];
  for (int k = 0; k < SQ_RTC_D; ++k) l = l + oc[k] * oc[k];
  const double disc = b * b - (l - p[SQ_RTC_D] * p[SQ_RTC_D]);
  if (disc < 0.0) return false;
  const double sq = __builtin_sqrt(disc);
  *t0 = -b - sq;
  *t1 = -b + sq;
  return true;
}
__device__ bool sq_narrow64(const double* p, const double* o, const double* d, double max_toi, double* toi) {
  double t0, t1;
  if (!shell64_roots(p, o, d, &t0, &t1)) return false;
  const double t = t0 >= 0.0 ? t0 : t1;
  if (t < 0.0 || t > max_toi) return false;
  *toi = t;
  return true;
}
template <class Out>
__device__ void sq_narrow64_crossings(const double* p, const double* o, const double* d, double max_toi, Out& out) {
  double t0, t1;
  if (!shell64_roots(p, o, d, &t0, &t1)) return;
  out(t0);
  out(t1);
}
"""


def crossings64_addr(lib):
    """Address of the crossings entry point of a compile_cpu64 library (0: none)."""
    return ctypes.cast(lib.sq_crossings64_c, ctypes.c_void_p).value if hasattr(lib, "sq_crossings64_c") else 0
