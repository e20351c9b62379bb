// rtc.hip -- consumer narrow phases, compiled at run time (hipRTC) into the
// nearest-hit walk.  The device analogue of GpuGeometry::narrow_wgsl
// (crates/spatial-query/src/gpu/mod.rs:86-99,261), which splices a consumer's
// WGSL narrow test into the reference's traversal shader: here the consumer's
// HIP source is compiled together with traverse.h and pool.h (embedded in
// the library) and the persistent pooled nearest-hit kernel, for gfx950.
//
// The consumer defines
//   __device__ bool sq_narrow(const float* prim, const float* origin,
//                             const float* dir, float max_toi, float* toi);
// (QueryGeometry::test_ray, geometry.rs:33-43: a hit at toi in [0, max_toi]
// or none) and, optionally, every crossing of the ray with the primitive
// (QueryGeometry::test_ray_crossings, geometry.rs:45-63), announced by
//   #define SQ_NARROW_CROSSINGS 1
//   template <class Out> __device__ void sq_narrow_crossings(const float* prim,
//       const float* origin, const float* dir, float max_toi, Out& out);
// calling out(toi) once per crossing, in any order.  Without it, crossings
// are the single sq_narrow hit (the reference's default).  Optionally too, a
// swept-box contact (QueryGeometry::test_shape_cast, geometry.rs:65-77) and an
// exact overlap test (test_overlap, geometry.rs:79-87):
//   #define SQ_NARROW_CAST 1
//   __device__ bool sq_narrow_cast(const float* prim, const float* origin,
//       const float* dir, const float* probe_min, const float* probe_max,
//       float max_toi, float* toi);
//   #define SQ_NARROW_OVERLAP 1
//   __device__ bool sq_narrow_overlap(const float* prim, const float* region_min,
//       const float* region_max);
// Without them a cast finds no contact and overlap is the box test (the
// trait's defaults).  The same module holds the all-hits, crossings, cast and
// overlap walks (pool_walk_with, pool_collect_with).  The tree is built from the consumer's world boxes
// (QueryGeometry::world_aabb); each leaf-ordered primitive record is its box
// followed by its prim_floats floats, of which sq_narrow sees the latter.  A
// hit counts as the reference counts it (in [0, max_toi]); nearest queries walk
// the reference's order (traverse.h ORDER_REF), so the answer is the
// reference's for any consumer, hits outside their own box included.  Under
// SQ_HITS_GUARDED, as for the built-in SPHERE and TRIANGLE kinds, a hit is
// kept only when its toi lies in the slab interval of the primitive's own box
// and the walk is near-first.
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include <map>
#include <mutex>
#include <tuple>

#include "engine_internal.h"
#include "pool.h"
#include "rtc_sources.inc"    // kTraverseSrc, kPoolSrc, kWalk64Src: traverse.h, pool.h, walk64.h as string literals (built by the Makefile)

namespace {

constexpr int KIND_CONSUMER = 3;

const char* kKernelSrc = R"SQRTC(
namespace sqd {
// GUARD keeps a hit only inside the slab interval of the primitive's box.
template <int D, bool GUARD> struct ConsumerPolicy {
  static constexpr bool INSIDE = GUARD;
  RayD<D> r;
  SQ_HD void setup(const float* src) { ray_setup<D>(r, src); }
  SQ_HD bool root(const float* root, float limit) const {
    float te, tx; int ax;
    return slab_exact<D>(r, root, root + D, limit, te, tx, ax);
  }
  SQ_HD void pair(const float* nd, float limit, bool& ha, bool& hb, float& ta, float& tb) const {
    slab_pair<D>(r, nd, limit, ha, hb, ta, tb);
  }
  __device__ bool narrow(const float* p, float limit, float& t) const {
    float tt;
    if (!sq_narrow(p + 2 * D, r.o, r.d, limit, &tt)) return false;
    if (!(tt >= 0.0f && tt <= limit)) return false;
    if constexpr (GUARD) {
      float te, tx; int ax;
      if (!slab_exact<D>(r, p, p + D, limit, te, tx, ax)) return false;
      if (!(te <= tt && tt <= tx)) return false;
    }
    t = tt;
    return true;
  }
};
}  // namespace sqd

// Persistent pooled walk (as the built-in kinds' k_nearest_pool): waves take
// chunks of `chunk` rays from counter[0] (zero at the start, left zero by
// pool_retire), refill lanes as rays finish, pause descents below pause_below
// lanes.  GUARD: the guarded, near-first walk; else the reference's order.
namespace sqd {
template <bool GUARD>
__device__ __forceinline__ void consumer_nearest(SceneView s, const float* rays, unsigned n, float max_toi, int* out_leaf,
                                                 float* out_toi, unsigned chunk, unsigned* counter, unsigned S,
                                                 SpillEntry* spill, unsigned spill_stride, unsigned pause_below) {
  extern __shared__ unsigned lds[];
  const unsigned B = blockDim.x, tid = threadIdx.x;
  const unsigned gtid = blockIdx.x * B + tid;
  sqd::RingStack st{lds + tid, reinterpret_cast<float*>(lds + S * B) + tid, B, S, spill + gtid, spill_stride};
  const unsigned long long waves = (unsigned long long)gridDim.x * (B >> 6);
  const unsigned long long home = (unsigned long long)blockIdx.x * (B >> 6) + (tid >> 6);
  const sqd::NearestOut out{out_leaf, out_toi, nullptr, nullptr, nullptr, nullptr};
  const sqd::Filter f{nullptr, ~0ull, -1};
  sqd::pool_walk_with<SQ_RTC_D, sqd::KIND_AABB, false, false, false, sqd::ConsumerPolicy<SQ_RTC_D, GUARD>, 2 * SQ_RTC_D,
                      2 * SQ_RTC_D + SQ_RTC_PF, sqd::Chunks, false, !GUARD>(
      s, rays, sqd::Chunks{counter, n, chunk, home, waves}, max_toi, nullptr, f, out, st, nullptr, pause_below);
  sqd::pool_retire(counter);
}
}  // namespace sqd

extern "C" __global__ __launch_bounds__(256) void sq_consumer_nearest(sqd::SceneView s, const float* rays, unsigned n,
                                                                      float max_toi, int* out_leaf, float* out_toi,
                                                                      unsigned chunk, unsigned* counter, unsigned S,
                                                                      sqd::SpillEntry* spill, unsigned spill_stride,
                                                                      unsigned pause_below) {
  sqd::consumer_nearest<false>(s, rays, n, max_toi, out_leaf, out_toi, chunk, counter, S, spill, spill_stride, pause_below);
}
extern "C" __global__ __launch_bounds__(256) void sq_consumer_nearest_guarded(sqd::SceneView s, const float* rays, unsigned n,
                                                                              float max_toi, int* out_leaf, float* out_toi,
                                                                              unsigned chunk, unsigned* counter, unsigned S,
                                                                              sqd::SpillEntry* spill, unsigned spill_stride,
                                                                              unsigned pause_below) {
  sqd::consumer_nearest<true>(s, rays, n, max_toi, out_leaf, out_toi, chunk, counter, S, spill, spill_stride, pause_below);
}

namespace sqd {
// Per-query state of an all-hits (CROSS false) or crossings (CROSS true) walk
// over consumer primitives: every primitive of every leaf range the ray
// reaches within max_toi is tested (no best-toi pruning; the parts of a
// reference leaf split into a synthetic chain are all tested), and each hit or
// crossing in [0, max_toi] is appended with its (toi, leaf) key.
template <int D, bool CROSS> struct ConsumerCollect {
  ConsumerPolicy<D, false> pol;
  float lim;
  SQ_HD bool begin(const SceneView& s, const float* q, unsigned i, float max_toi, const float*) {
    pol.setup(q + (size_t)i * 2 * D);
    lim = max_toi;
    return pol.root(s.root, lim);
  }
  SQ_HD void pair(const float* nd, bool& ha, bool& hb) const {
    if (synthetic_pair<D>(nd)) { ha = hb = true; return; }
    float ta, tb;
    pol.pair(nd, lim, ha, hb, ta, tb);
  }
  template <class Emit>
  SQ_HD void leaf(const SceneView& s, const Filter&, unsigned first, unsigned c, Emit&& emit) const {
    for (unsigned j = 0; j < c; ++j) {
      const unsigned pos = first + j;
      const float* p = s.prims + (size_t)pos * (2 * D + SQ_RTC_PF);
      const unsigned leaf = s.ids[pos];
#if defined(SQ_NARROW_CROSSINGS)
      if constexpr (CROSS) {
        auto out = [&](float t) {
          if (t >= 0.0f && t <= lim) emit(hit_key(t, leaf));
        };
        sq_narrow_crossings(p + 2 * D, pol.r.o, pol.r.d, lim, out);
        continue;
      }
#endif
      float t;
      if (pol.narrow(p, lim, t)) emit(hit_key(t, leaf));
    }
  }
};

// Shape casts through the consumer (QueryGeometry::test_shape_cast,
// geometry.rs:65-77): queries [4D] = origin, dir, probe box min, max; the
// probe-grown slab prunes (ShapeCast::broadphase, shapecast.rs:70-73) and a
// contact counts when its toi is in [0, limit] (with GUARD, also inside the
// grown slab interval of the primitive's own box).  Without SQ_NARROW_CAST
// there is no contact (the trait's default).
template <int D, bool GUARD> struct ConsumerCastPolicy : CastPolicy<D, KIND_AABB> {
  static constexpr bool INSIDE = GUARD;
  __device__ bool narrow(const float* p, float limit, float& t) const {
#if defined(SQ_NARROW_CAST)
    float tt;
    if (!sq_narrow_cast(p + 2 * D, this->r.o, this->r.d, this->pmin, this->pmax, limit, &tt)) return false;
    if (!(tt >= 0.0f && tt <= limit)) return false;
    if constexpr (GUARD) {
      float te, tx;
      if (!this->grown_slab(p, p + D, limit, te, tx)) return false;
      if (!(te <= tt && tt <= tx)) return false;
    }
    t = tt;
    return true;
#else
    (void)p; (void)limit; (void)t;
    return false;
#endif
  }
};

// All shape-cast contacts (shapecast_all, bvh.rs:381-414) per cast.
template <int D> struct ConsumerCastAll {
  ConsumerCastPolicy<D, false> pol;
  float lim;
  SQ_HD bool begin(const SceneView& s, const float* q, unsigned i, float max_toi, const float*) {
    pol.setup(q + (size_t)i * 4 * D);
    lim = max_toi;
    return pol.root(s.root, lim);
  }
  SQ_HD void pair(const float* nd, bool& ha, bool& hb) const {
    if (synthetic_pair<D>(nd)) { ha = hb = true; return; }
    float ta, tb;
    pol.pair(nd, lim, ha, hb, ta, tb);
  }
  template <class Emit>
  SQ_HD void leaf(const SceneView& s, const Filter&, unsigned first, unsigned c, Emit&& emit) const {
    for (unsigned j = 0; j < c; ++j) {
      const unsigned pos = first + j;
      float t;
      if (pol.narrow(s.prims + (size_t)pos * (2 * D + SQ_RTC_PF), lim, t)) emit(hit_key(t, s.ids[pos]));
    }
  }
};

// Overlap (Bvh::overlap, bvh.rs:418-456) per region [2D]: every primitive
// whose world box meets the region and that passes the consumer's
// sq_narrow_overlap (QueryGeometry::test_overlap, geometry.rs:79-87; without
// SQ_NARROW_OVERLAP the box test alone, the trait's default), keyed by leaf.
template <int D> struct ConsumerOverlap {
  float mn[D], mx[D];
  SQ_HD bool begin(const SceneView& s, const float* q, unsigned i, float, const float*) {
    const float* p = q + (size_t)i * 2 * D;
#pragma unroll
    for (int k = 0; k < D; ++k) { mn[k] = p[k]; mx[k] = p[D + k]; }
    return box_overlap<D>(s.root, s.root + D, mn, mx);
  }
  SQ_HD void pair(const float* nd, bool& ha, bool& hb) const {
    float amn[D], amx[D], bmn[D], bmx[D];
#pragma unroll
    for (int k = 0; k < D; ++k) { amn[k] = nd[4 * k]; amx[k] = nd[4 * k + 1]; bmn[k] = nd[4 * k + 2]; bmx[k] = nd[4 * k + 3]; }
    ha = box_overlap<D>(amn, amx, mn, mx);
    hb = box_overlap<D>(bmn, bmx, mn, mx);
  }
  template <class Emit>
  SQ_HD void leaf(const SceneView& s, const Filter&, unsigned first, unsigned c, Emit&& emit) const {
    for (unsigned j = 0; j < c; ++j) {
      const unsigned pos = first + j;
      const float* p = s.prims + (size_t)pos * (2 * D + SQ_RTC_PF);
      if (!box_overlap<D>(p, p + D, mn, mx)) continue;
#if defined(SQ_NARROW_OVERLAP)
      if (!sq_narrow_overlap(p + 2 * D, mn, mx)) continue;
#endif
      emit((unsigned long long)s.ids[pos]);
    }
  }
};

template <class CQ>
__device__ __forceinline__ void consumer_collect(const SceneView& s, const float* rays, unsigned n, float max_toi,
                                                 unsigned chunk, unsigned* counter, unsigned S, SpillEntry* spill,
                                                 unsigned spill_stride, unsigned pause_below, const StageOut& o) {
  extern __shared__ unsigned lds[];
  const unsigned B = blockDim.x, tid = threadIdx.x;
  const unsigned gtid = blockIdx.x * B + tid;
  RingStack st{lds + tid, reinterpret_cast<float*>(lds + S * B) + tid, B, S, spill + gtid, spill_stride};
  volatile unsigned* slot = lds + 2 * S * B + 2 * (tid >> 6);
  if ((tid & 63u) == 0) {
    slot[0] = NO_BLOCK;
    slot[1] = 0;
  }
  const unsigned long long waves = (unsigned long long)gridDim.x * (B >> 6);
  const unsigned long long home = (unsigned long long)blockIdx.x * (B >> 6) + (tid >> 6);
  const Filter f{nullptr, ~0ull, -1};
  pool_collect_with<SQ_RTC_D, CQ, false>(s, rays, Chunks{counter, n, chunk, home, waves},
                                                                        max_toi, nullptr, f, st, StageAppend{o, slot},
                                                                        pause_below, nullptr, nullptr);
}
}  // namespace sqd

// All hits / all crossings of each ray (pooled collect): waves take chunks of
// `chunk` rays from counter[0] (cleared by the host with the staging words:
// counter[1] blocks claimed, counter[2] overflow); entries go to the staging
// blocks of o and per-ray counts to o.counts.  LDS: the ring stacks, then two
// words per wave.
extern "C" __global__ __launch_bounds__(256) void sq_consumer_all(sqd::SceneView s, const float* rays, unsigned n,
                                                                  float max_toi, unsigned chunk, unsigned* counter,
                                                                  unsigned S, sqd::SpillEntry* spill, unsigned spill_stride,
                                                                  unsigned pause_below, sqd::StageOut o) {
  sqd::consumer_collect<sqd::ConsumerCollect<SQ_RTC_D, false>>(s, rays, n, max_toi, chunk, counter, S, spill, spill_stride,
                                                               pause_below, o);
}
extern "C" __global__ __launch_bounds__(256) void sq_consumer_crossings(sqd::SceneView s, const float* rays, unsigned n,
                                                                        float max_toi, unsigned chunk, unsigned* counter,
                                                                        unsigned S, sqd::SpillEntry* spill,
                                                                        unsigned spill_stride, unsigned pause_below,
                                                                        sqd::StageOut o) {
  sqd::consumer_collect<sqd::ConsumerCollect<SQ_RTC_D, true>>(s, rays, n, max_toi, chunk, counter, S, spill, spill_stride,
                                                              pause_below, o);
}
extern "C" __global__ __launch_bounds__(256) void sq_consumer_cast_all(sqd::SceneView s, const float* casts, unsigned n,
                                                                       float max_toi, unsigned chunk, unsigned* counter,
                                                                       unsigned S, sqd::SpillEntry* spill,
                                                                       unsigned spill_stride, unsigned pause_below,
                                                                       sqd::StageOut o) {
  sqd::consumer_collect<sqd::ConsumerCastAll<SQ_RTC_D>>(s, casts, n, max_toi, chunk, counter, S, spill, spill_stride,
                                                        pause_below, o);
}
extern "C" __global__ __launch_bounds__(256) void sq_consumer_overlap(sqd::SceneView s, const float* regions, unsigned n,
                                                                      float max_toi, unsigned chunk, unsigned* counter,
                                                                      unsigned S, sqd::SpillEntry* spill,
                                                                      unsigned spill_stride, unsigned pause_below,
                                                                      sqd::StageOut o) {
  sqd::consumer_collect<sqd::ConsumerOverlap<SQ_RTC_D>>(s, regions, n, max_toi, chunk, counter, S, spill, spill_stride,
                                                        pause_below, o);
}

// Nearest shape-cast contact per cast [4D] (shapecast_nearest,
// bvh.rs:331-378), the pooled walk with the consumer's cast policy (GUARD as
// consumer_nearest).
namespace sqd {
template <bool GUARD>
__device__ __forceinline__ void consumer_cast_nearest(SceneView s, const float* casts, unsigned n, float max_toi,
                                                      int* out_leaf, float* out_toi, unsigned chunk, unsigned* counter,
                                                      unsigned S, SpillEntry* spill, unsigned spill_stride,
                                                      unsigned pause_below) {
  extern __shared__ unsigned lds[];
  const unsigned B = blockDim.x, tid = threadIdx.x;
  const unsigned gtid = blockIdx.x * B + tid;
  sqd::RingStack st{lds + tid, reinterpret_cast<float*>(lds + S * B) + tid, B, S, spill + gtid, spill_stride};
  const unsigned long long waves = (unsigned long long)gridDim.x * (B >> 6);
  const unsigned long long home = (unsigned long long)blockIdx.x * (B >> 6) + (tid >> 6);
  const sqd::NearestOut out{out_leaf, out_toi, nullptr, nullptr, nullptr, nullptr};
  const sqd::Filter f{nullptr, ~0ull, -1};
  sqd::pool_walk_with<SQ_RTC_D, sqd::KIND_AABB, false, false, true, sqd::ConsumerCastPolicy<SQ_RTC_D, GUARD>, 4 * SQ_RTC_D,
                      2 * SQ_RTC_D + SQ_RTC_PF, sqd::Chunks, false, !GUARD>(
      s, casts, sqd::Chunks{counter, n, chunk, home, waves}, max_toi, nullptr, f, out, st, nullptr, pause_below);
  sqd::pool_retire(counter);
}
}  // namespace sqd

extern "C" __global__ __launch_bounds__(256) void sq_consumer_cast_nearest(sqd::SceneView s, const float* casts, unsigned n,
                                                                           float max_toi, int* out_leaf, float* out_toi,
                                                                           unsigned chunk, unsigned* counter, unsigned S,
                                                                           sqd::SpillEntry* spill, unsigned spill_stride,
                                                                           unsigned pause_below) {
  sqd::consumer_cast_nearest<false>(s, casts, n, max_toi, out_leaf, out_toi, chunk, counter, S, spill, spill_stride, pause_below);
}
extern "C" __global__ __launch_bounds__(256) void sq_consumer_cast_nearest_guarded(sqd::SceneView s, const float* casts,
                                                                                   unsigned n, float max_toi, int* out_leaf,
                                                                                   float* out_toi, unsigned chunk,
                                                                                   unsigned* counter, unsigned S,
                                                                                   sqd::SpillEntry* spill,
                                                                                   unsigned spill_stride,
                                                                                   unsigned pause_below) {
  sqd::consumer_cast_nearest<true>(s, casts, n, max_toi, out_leaf, out_toi, chunk, counter, S, spill, spill_stride, pause_below);
}
)SQRTC";

}  // namespace

struct sq_narrow {
  sq_ctx* ctx = nullptr;
  int dim = 0;
  uint32_t pf = 0;
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;          // nearest hit
  hipFunction_t fn_g = nullptr;        // nearest hit, SQ_HITS_GUARDED
  hipFunction_t fn_cast_g = nullptr;   // nearest shape-cast contact, SQ_HITS_GUARDED
  hipFunction_t fn_all = nullptr;      // all hits
  hipFunction_t fn_cross = nullptr;    // all crossings
  hipFunction_t fn_cast = nullptr;     // nearest shape-cast contact
  hipFunction_t fn_cast_all = nullptr; // all shape-cast contacts
  hipFunction_t fn_overlap = nullptr;  // overlap
  // resident blocks per CU by (kernel, block, LDS), queried once each
  std::mutex occ_mu;
  std::map<std::tuple<hipFunction_t, uint32_t, size_t>, int> occ;
};

// Resident blocks per CU of one of the module's kernels.
static int module_blocks_per_cu(sq_narrow* nw, hipFunction_t fn, uint32_t block, size_t lds, int* out) {
  const auto key = std::make_tuple(fn, block, lds);
  {
    std::lock_guard<std::mutex> g(nw->occ_mu);
    auto it = nw->occ.find(key);
    if (it != nw->occ.end()) { *out = it->second; return SQ_OK; }
  }
  int per_cu = 0;
  HIP_TRY(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)block, lds));
  std::lock_guard<std::mutex> g(nw->occ_mu);
  nw->occ[key] = per_cu;
  *out = per_cu;
  return SQ_OK;
}

// Compiles a consumer's source for gfx950 (hipRTC) after `#include <header>`
// (one of the embedded traverse.h, pool.h, walk64.h) and before `kernels`,
// with SQ_RTC_D = dim and SQ_RTC_PF = pf, and loads the module on the
// context's device.  A source that does not compile is SQ_ERR_INVALID with the
// compiler's log.
int rtc_load_module(sq_ctx* ctx, const char* header, const char* source, const char* kernels, int dim, uint32_t pf,
                    hipModule_t* out) {
  HIP_TRY(hipSetDevice(ctx->device));
  const std::string src = std::string("#include \"") + header + "\"\n#line 1 \"consumer\"\n" + source + "\n" + kernels;
  const char* headers[] = {kTraverseSrc, kPoolSrc, kWalk64Src};
  const char* names[] = {"traverse.h", "pool.h", "walk64.h"};
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "consumer_narrow.hip", 3, headers, names) != HIPRTC_SUCCESS)
    return fail(SQ_ERR_HIP, "hiprtcCreateProgram failed");
  const std::string dd = "-DSQ_RTC_D=" + std::to_string(dim), pp = "-DSQ_RTC_PF=" + std::to_string(pf);
  // the engine's own flags: no contraction, f32 denormals, no packed FP32
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++20", "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", dd.c_str(), pp.c_str()};
  const hiprtcResult cr = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  size_t log_n = 0;
  hiprtcGetProgramLogSize(prog, &log_n);
  std::string log(log_n, '\0');
  if (log_n) hiprtcGetProgramLog(prog, &log[0]);
  if (cr != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return fail(SQ_ERR_INVALID, "consumer narrow phase does not compile:\n" + log);
  }
  size_t code_n = 0;
  hiprtcGetCodeSize(prog, &code_n);
  std::vector<char> code(code_n);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  HIP_TRY(hipModuleLoadData(out, code.data()));
  return SQ_OK;
}

extern "C" {

int sq_narrow_create(sq_ctx* ctx, int dim, const char* source, uint32_t prim_floats, sq_narrow** out) {
  if (!ctx || !source || !out) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (dim < 2 || dim > 6) return fail(SQ_ERR_UNSUPPORTED, "consumer narrow phases take dim 2..6");
  if (prim_floats > 256) return fail(SQ_ERR_INVALID, "at most 256 floats per primitive");
  auto nw = std::make_unique<sq_narrow>();
  nw->ctx = ctx;
  nw->dim = dim;
  nw->pf = prim_floats;
  if (int rc = rtc_load_module(ctx, "pool.h", source, kKernelSrc, dim, prim_floats, &nw->mod)) return rc;
  if (hipModuleGetFunction(&nw->fn, nw->mod, "sq_consumer_nearest") != hipSuccess ||
      hipModuleGetFunction(&nw->fn_all, nw->mod, "sq_consumer_all") != hipSuccess ||
      hipModuleGetFunction(&nw->fn_cross, nw->mod, "sq_consumer_crossings") != hipSuccess ||
      hipModuleGetFunction(&nw->fn_cast, nw->mod, "sq_consumer_cast_nearest") != hipSuccess ||
      hipModuleGetFunction(&nw->fn_g, nw->mod, "sq_consumer_nearest_guarded") != hipSuccess ||
      hipModuleGetFunction(&nw->fn_cast_g, nw->mod, "sq_consumer_cast_nearest_guarded") != hipSuccess ||
      hipModuleGetFunction(&nw->fn_cast_all, nw->mod, "sq_consumer_cast_all") != hipSuccess ||
      hipModuleGetFunction(&nw->fn_overlap, nw->mod, "sq_consumer_overlap") != hipSuccess) {
    (void)hipModuleUnload(nw->mod);
    return fail(SQ_ERR_HIP, "consumer kernel not found in the compiled module");
  }
  ctx_retain(ctx);
  *out = nw.release();
  return SQ_OK;
}

int sq_narrow_destroy(sq_narrow* nw) {
  if (!nw) return SQ_OK;
  sq_ctx* c = nw->ctx;
  (void)hipStreamSynchronize(c->stream);
  if (nw->mod) (void)hipModuleUnload(nw->mod);
  delete nw;
  ctx_release(c);
  return SQ_OK;
}

int sq_bvh_build_consumer(sq_narrow* nw, const float* boxes, const float* prims, uint64_t n, sq_bvh** out) {
  if (!nw || !out || (n && (!boxes || !prims))) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (n > sqh::LEAF_FIRST_MASK) return fail(SQ_ERR_INVALID, "at most 2^27 primitives per tree");
  sq_ctx* ctx = nw->ctx;
  const int D = nw->dim;
  auto b = std::make_unique<sq_bvh>();
  b->dim = D;
  b->kind = KIND_CONSUMER;
  b->n_prims = n;
  b->consumer_pf = nw->pf;
  sqh::build_tree(D, boxes, (uint32_t)n, b->tree);
  try {
    sqh::pack_tree(b->tree, b->packed);
  } catch (const std::exception& e) {
    return fail(SQ_ERR_INVALID, e.what());
  }
  // leaf-ordered records: the world box, then the consumer's floats
  const size_t rs = 2 * (size_t)D + nw->pf;
  std::vector<float> dp(std::max<size_t>(n, 1) * rs, 0.0f);
  for (size_t k = 0; k < n; ++k) {
    const uint32_t id = b->tree.prim_indices[k];
    std::memcpy(dp.data() + k * rs, boxes + (size_t)id * 2 * D, 2 * D * sizeof(float));
    if (nw->pf) std::memcpy(dp.data() + k * rs + 2 * D, prims + (size_t)id * nw->pf, nw->pf * sizeof(float));
  }
  hipStream_t s = ctx->stream;
  HIP_TRY(hipSetDevice(ctx->device));
  int rc;
  if ((rc = upload(b->nodes, b->packed.nodes.data(), b->packed.nodes.size() * 4, s))) return rc;
  if ((rc = upload(b->prims, dp.data(), dp.size() * 4, s))) return rc;
  if ((rc = upload(b->ids, b->tree.prim_indices.data(), b->tree.prim_indices.size() * 4, s))) return rc;
  if ((rc = upload(b->root, b->packed.root.data(), b->packed.root.size() * 4, s))) return rc;
  HIP_TRY(hipStreamSynchronize(s));
  tree_stats_of(b->tree, b->stats);
  b->ctx = ctx;
  ctx_retain(ctx);
  *out = b.release();
  return SQ_OK;
}

}  // extern "C"

// The tree was built for this narrow phase (same dimension, record size and
// context) and holds consumer records.
static int check_pair(const sq_narrow* nw, const sq_bvh* b) {
  if (b->kind != KIND_CONSUMER || b->dim != nw->dim || b->consumer_pf != nw->pf || b->ctx != nw->ctx)
    return fail(SQ_ERR_INVALID, "the tree was not built for this narrow phase");
  return SQ_OK;
}

// One pooled nearest launch of module kernel fn (sq_consumer_nearest or
// sq_consumer_cast_nearest, queries of qs floats), as nearest_impl launches
// the built-in kinds: chunks of 64 * per queries, a grid of the resident
// waves, ring stacks spilling.
static int consumer_nearest(sq_narrow* nw, sq_bvh* b, hipFunction_t fn, const float* q, uint64_t n, float max_toi,
                            int32_t* out_leaf, float* out_toi) {
  if (!nw || !b || (n && (!q || !out_leaf || !out_toi))) return fail(SQ_ERR_INVALID, "NULL argument");
  int rc;
  if ((rc = check_pair(nw, b))) return rc;
  if (n > 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "at most 2^32-1 queries per call");
  if (!n) return SQ_OK;
  sq_ctx* c = b->ctx;
  if (!b->n_prims) return fill_miss(c, n, b->dim, NearestOut{out_leaf, out_toi, nullptr, nullptr, nullptr, nullptr});
  const uint32_t B = (uint32_t)c->block;
  uint32_t S = (uint32_t)c->lds_stack;
  const uint32_t per = pooled_per_lane(c, n);
  const uint64_t chunk64 = 64ull * per;
  const uint64_t chunks = (n + chunk64 - 1) / chunk64;
  const uint32_t stack_n = b->packed.max_stack;
  const uint32_t spill_rows = stack_n > S ? stack_n - S : 0;
  const size_t lds = (size_t)S * B * 8;
  int per_cu = 0;
  if ((rc = module_blocks_per_cu(nw, fn, B, lds, &per_cu))) return rc;
  const uint64_t wpb = B / 64;
  const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)c->cus * std::max(1, per_cu), (chunks + wpb - 1) / wpb));
  uint32_t lanes = (uint32_t)(grid * B);
  if ((rc = c->spill.ensure((size_t)spill_rows * lanes * sizeof(SpillEntry) + 64))) return rc;
  SceneView s = view_of(b);
  const float* qp = q;
  uint32_t nn = (uint32_t)n, chunk = (uint32_t)chunk64, pause = (uint32_t)c->descend_min;
  int32_t* lp = out_leaf;
  float* tp = out_toi;
  uint32_t* counter = c->pool_ctr.as<uint32_t>();
  SpillEntry* spill = c->spill.as<SpillEntry>();
  void* args[] = {&s, &qp, &nn, &max_toi, &lp, &tp, &chunk, &counter, &S, &spill, &lanes, &pause};
  HIP_TRY(hipModuleLaunchKernel(fn, (unsigned)grid, 1, 1, B, 1, 1, (unsigned)lds, c->stream, args, nullptr));
  return SQ_OK;
}

// Variable-length results through module kernel fn (all hits, crossings,
// cast contacts or overlaps): the pooled collect of the built-in kinds
// (collect_impl without sub-walks) and its staged CSR pipeline.
static int consumer_collect(sq_narrow* nw, sq_bvh* b, hipFunction_t fn, const float* q, uint64_t n, float max_toi,
                            bool has_t, uint64_t* offsets, int32_t* out_leaf, float* out_toi, uint64_t capacity,
                            uint64_t* total) {
  if (!nw || !b || !offsets || !total || (n && !q)) return fail(SQ_ERR_INVALID, "NULL argument");
  int rc;
  if ((rc = check_pair(nw, b))) return rc;
  if (n > 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "at most 2^32-1 queries per call");
  sq_ctx* c = b->ctx;
  *total = 0;
  if (!n || !b->n_prims) {
    HIP_TRY(hipMemsetAsync(offsets, 0, (n + 1) * 8, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SQ_OK;
  }
  if ((rc = c->counts.ensure(n * 4))) return rc;
  if ((rc = c->counter.ensure(256))) return rc;
  const uint32_t B = (uint32_t)c->block;
  uint32_t S = (uint32_t)c->lds_stack;
  const uint32_t per = pooled_per_lane(c, n);
  const uint64_t chunk64 = 64ull * per;
  const uint64_t chunks = (n + chunk64 - 1) / chunk64;
  const uint32_t stack_n = b->packed.max_stack;
  const uint32_t spill_rows = stack_n > S ? stack_n - S : 0;
  const size_t lds = (size_t)S * B * 8 + (size_t)(B / 64) * 8;
  int per_cu = 0;
  if ((rc = module_blocks_per_cu(nw, fn, B, lds, &per_cu))) return rc;
  const uint64_t wpb = B / 64;
  const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)c->cus * std::max(1, per_cu), (chunks + wpb - 1) / wpb));
  uint32_t lanes = (uint32_t)(grid * B);
  auto launch = [&](uint64_t want, uint64_t, StagedWalk& walk) -> int {
    int r;
    if ((r = c->spill.ensure((size_t)spill_rows * lanes * sizeof(SpillEntry) + 64))) return r;
    // retired blocks hold over STAGE_BLK - 64 entries; each wave has one more open
    const uint64_t blocks = std::min<uint64_t>(want / (STAGE_BLK - 64) + 1 + grid * wpb, 0xFFFFFFF0ull);
    if ((r = c->stage.ensure(blocks * STAGE_BLK * sizeof(StageEntry)))) return r;
    if ((r = c->stage_used.ensure(blocks * 4))) return r;
    const uint64_t max_blocks =
        std::min<uint64_t>({c->stage.bytes / (STAGE_BLK * sizeof(StageEntry)), c->stage_used.bytes / 4, 0xFFFFFFF0ull});
    uint32_t* counter = c->counter.as<uint32_t>();
    HIP_TRY(hipMemsetAsync(counter, 0, 64, c->stream));
    StageOut o{c->stage.as<StageEntry>(), c->stage_used.as<uint32_t>(), counter + 1, counter + 2, (uint32_t)max_blocks,
               c->counts.as<uint32_t>()};
    SceneView s = view_of(b);
    const float* qp = q;
    uint32_t nn = (uint32_t)n, chunk = (uint32_t)chunk64, pause = (uint32_t)c->descend_min;
    SpillEntry* spill = c->spill.as<SpillEntry>();
    void* args[] = {&s, &qp, &nn, &max_toi, &chunk, &counter, &S, &spill, &lanes, &pause, &o};
    HIP_TRY(hipModuleLaunchKernel(fn, (unsigned)grid, 1, 1, B, 1, 1, (unsigned)lds, c->stream, args, nullptr));
    walk = StagedWalk{nullptr, nullptr};
    return SQ_OK;
  };
  return csr_run_staged(c, n, offsets, out_leaf, out_toi, has_t, capacity, total, launch);
}

extern "C" {

int sq_raycast_nearest_consumer(sq_narrow* nw, sq_bvh* b, const float* rays, uint64_t n, float max_toi, int32_t* out_leaf,
                                float* out_toi) {
  return consumer_nearest(nw, b, nw ? (nw->ctx->hit_mode == SQ_HITS_GUARDED ? nw->fn_g : nw->fn) : nullptr, rays, n, max_toi,
                          out_leaf, out_toi);
}

int sq_raycast_all_consumer(sq_narrow* nw, sq_bvh* b, const float* rays, uint64_t n, float max_toi, int crossings,
                            uint64_t* offsets, int32_t* out_leaf, float* out_toi, uint64_t capacity, uint64_t* total) {
  if (!nw) return fail(SQ_ERR_INVALID, "NULL argument");
  return consumer_collect(nw, b, crossings ? nw->fn_cross : nw->fn_all, rays, n, max_toi, true, offsets, out_leaf, out_toi,
                          capacity, total);
}

int sq_shapecast_nearest_consumer(sq_narrow* nw, sq_bvh* b, const float* casts, uint64_t n, float max_toi, int32_t* out_leaf,
                                  float* out_toi) {
  return consumer_nearest(nw, b, nw ? (nw->ctx->hit_mode == SQ_HITS_GUARDED ? nw->fn_cast_g : nw->fn_cast) : nullptr, casts, n,
                          max_toi, out_leaf, out_toi);
}

int sq_shapecast_all_consumer(sq_narrow* nw, sq_bvh* b, const float* casts, uint64_t n, float max_toi, uint64_t* offsets,
                              int32_t* out_leaf, float* out_toi, uint64_t capacity, uint64_t* total) {
  if (!nw) return fail(SQ_ERR_INVALID, "NULL argument");
  return consumer_collect(nw, b, nw->fn_cast_all, casts, n, max_toi, true, offsets, out_leaf, out_toi, capacity, total);
}

int sq_overlap_consumer(sq_narrow* nw, sq_bvh* b, const float* regions, uint64_t n, uint64_t* offsets, int32_t* out_leaf,
                        uint64_t capacity, uint64_t* total) {
  if (!nw) return fail(SQ_ERR_INVALID, "NULL argument");
  return consumer_collect(nw, b, nw->fn_overlap, regions, n, 0.0f, false, offsets, out_leaf, nullptr, capacity, total);
}

}  // extern "C"
