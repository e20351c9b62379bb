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


def compile_cpu(source: str, dim: int, pf: int, workdir: str):
    """The consumer source built for the CPU (g++ -O2 -ffp-contract=off) as a
    shared object; returns (library, address of its C entry point)."""
    code = (f"#define __device__\n#define SQ_RTC_D {dim}\n#define SQ_RTC_PF {pf}\n" + source +
            '\nextern "C" int sq_narrow_c(const float* p, const float* o, const float* d, float m, float* t) '
            "{ return sq_narrow(p, o, d, m, t) ? 1 : 0; }\n")
    # one file name per source and shape: dlopen returns an already loaded
    # library for a path it has seen
    tag = hashlib.sha1(code.encode()).hexdigest()[:12]
    src = os.path.join(workdir, f"narrow_{tag}.cpp")
    so = os.path.join(workdir, f"narrow_{tag}.so")
    with open(src, "w") as f:
        f.write(code)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-o", so, src], check=True)
    lib = ctypes.CDLL(so)
    return lib, ctypes.cast(lib.sq_narrow_c, ctypes.c_void_p).value
