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
// or none).  The tree is built from the consumer's world boxes
// (QueryGeometry::world_aabb); each leaf-ordered primitive record is its box
// followed by its prim_floats floats, of which sq_narrow sees the latter.  As
// for the built-in SPHERE and TRIANGLE kinds, a hit is kept only when its toi
// lies in the slab interval of the primitive's own box, which makes the
// result the minimum under (toi, leaf) over the kept hits, whatever the order
// of the walk (the argument of traverse.h).
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "engine_internal.h"
#include "rtc_sources.inc"    // kTraverseSrc, kPoolSrc: traverse.h and pool.h as string literals (built by the Makefile)

namespace {

constexpr int KIND_CONSUMER = 3;

const char* kKernelSrc = R"SQRTC(
namespace sqd {
template <int D> struct ConsumerPolicy {
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
    float te, tx; int ax;
    if (!slab_exact<D>(r, p, p + D, limit, te, tx, ax)) return false;
    if (!(te <= tt && tt <= tx)) return false;
    t = tt;
    return true;
  }
};
}  // namespace sqd

// Persistent pooled walk (as the built-in kinds' k_nearest_pool): waves take
// chunks of `chunk` rays from counter[0] (zero at the start, left zero by
// pool_retire), refill lanes as rays finish, pause descents below pause_below
// lanes while the pool has rays left.
extern "C" __global__ __launch_bounds__(256) void sq_consumer_nearest(sqd::SceneView s, const float* rays, unsigned n,
                                                                      float max_toi, int* out_leaf, float* out_toi,
                                                                      unsigned chunk, unsigned* counter, unsigned S,
                                                                      sqd::SpillEntry* spill, unsigned spill_stride,
                                                                      unsigned pause_below) {
  extern __shared__ unsigned lds[];
  const unsigned B = blockDim.x, tid = threadIdx.x;
  const unsigned gtid = blockIdx.x * B + tid;
  sqd::RingStack st{lds + tid, reinterpret_cast<float*>(lds + S * B) + tid, B, S, spill + gtid, spill_stride};
  const unsigned long long waves = (unsigned long long)gridDim.x * (B >> 6);
  const unsigned long long home = (unsigned long long)blockIdx.x * (B >> 6) + (tid >> 6);
  const sqd::NearestOut out{out_leaf, out_toi, nullptr, nullptr, nullptr, nullptr};
  const sqd::Filter f{nullptr, ~0ull, -1};
  sqd::pool_walk_with<SQ_RTC_D, sqd::KIND_AABB, false, false, false, sqd::ConsumerPolicy<SQ_RTC_D>, 2 * SQ_RTC_D,
                      2 * SQ_RTC_D + SQ_RTC_PF>(s, rays, sqd::Chunks{counter, n, chunk, home, waves}, max_toi, nullptr, f,
                                                out, st, nullptr, pause_below);
  sqd::pool_retire(counter);
}
)SQRTC";

}  // namespace

struct sq_narrow {
  sq_ctx* ctx = nullptr;
  int dim = 0;
  uint32_t pf = 0;
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  // resident blocks per CU, cached for the last (block, LDS) pair asked
  uint32_t occ_block = 0;
  size_t occ_lds = 0;
  int occ_per_cu = -1;
};

extern "C" {

int sq_narrow_create(sq_ctx* ctx, int dim, const char* source, uint32_t prim_floats, sq_narrow** out) {
  if (!ctx || !source || !out) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (dim < 2 || dim > 6) return fail(SQ_ERR_UNSUPPORTED, "consumer narrow phases take dim 2..6");
  if (prim_floats > 256) return fail(SQ_ERR_INVALID, "at most 256 floats per primitive");
  HIP_TRY(hipSetDevice(ctx->device));
  const std::string src = std::string("#include \"pool.h\"\n#line 1 \"consumer\"\n") + source + "\n" + kKernelSrc;
  const char* headers[] = {kTraverseSrc, kPoolSrc};
  const char* names[] = {"traverse.h", "pool.h"};
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "consumer_narrow.hip", 2, headers, names) != HIPRTC_SUCCESS)
    return fail(SQ_ERR_HIP, "hiprtcCreateProgram failed");
  const std::string dd = "-DSQ_RTC_D=" + std::to_string(dim), pp = "-DSQ_RTC_PF=" + std::to_string(prim_floats);
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
  auto nw = std::make_unique<sq_narrow>();
  nw->ctx = ctx;
  nw->dim = dim;
  nw->pf = prim_floats;
  HIP_TRY(hipModuleLoadData(&nw->mod, code.data()));
  if (hipModuleGetFunction(&nw->fn, nw->mod, "sq_consumer_nearest") != hipSuccess) {
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

int sq_raycast_nearest_consumer(sq_narrow* nw, sq_bvh* b, const float* rays, uint64_t n, float max_toi, int32_t* out_leaf,
                                float* out_toi) {
  if (!nw || !b || (n && (!rays || !out_leaf || !out_toi))) return fail(SQ_ERR_INVALID, "NULL argument");
  if (b->kind != KIND_CONSUMER || b->dim != nw->dim || b->consumer_pf != nw->pf || b->ctx != nw->ctx)
    return fail(SQ_ERR_INVALID, "the tree was not built for this narrow phase");
  if (n > 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "at most 2^32-1 queries per call");
  if (!n) return SQ_OK;
  sq_ctx* c = b->ctx;
  if (!b->n_prims) return fill_miss(c, n, b->dim, NearestOut{out_leaf, out_toi, nullptr, nullptr, nullptr, nullptr});
  // the pooled launch of the built-in kinds (nearest_impl): chunks of
  // 64 * per rays, a grid of the resident waves, ring stacks spilling
  const uint32_t B = (uint32_t)c->block;
  uint32_t S = (uint32_t)c->lds_stack;
  const uint32_t per = pooled_per_lane(c, n);
  const uint64_t chunk64 = 64ull * per;
  const uint64_t chunks = (n + chunk64 - 1) / chunk64;
  const uint32_t stack_n = b->packed.max_stack;
  const uint32_t spill_rows = stack_n > S ? stack_n - S : 0;
  const size_t lds = (size_t)S * B * 8;
  int per_cu = 0;
  if (nw->occ_per_cu >= 0 && nw->occ_block == B && nw->occ_lds == lds) {
    per_cu = nw->occ_per_cu;
  } else {
    HIP_TRY(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, nw->fn, (int)B, lds));
    nw->occ_block = B;
    nw->occ_lds = lds;
    nw->occ_per_cu = per_cu;
  }
  const uint64_t wpb = B / 64;
  const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)c->cus * std::max(1, per_cu), (chunks + wpb - 1) / wpb));
  uint32_t lanes = (uint32_t)(grid * B);
  int rc;
  if ((rc = c->spill.ensure((size_t)spill_rows * lanes * sizeof(SpillEntry) + 64))) return rc;
  SceneView s = view_of(b);
  const float* rp = rays;
  uint32_t nn = (uint32_t)n, chunk = (uint32_t)chunk64, pause = (uint32_t)c->descend_min;
  int32_t* lp = out_leaf;
  float* tp = out_toi;
  uint32_t* counter = c->pool_ctr.as<uint32_t>();
  SpillEntry* spill = c->spill.as<SpillEntry>();
  void* args[] = {&s, &rp, &nn, &max_toi, &lp, &tp, &chunk, &counter, &S, &spill, &lanes, &pause};
  HIP_TRY(hipModuleLaunchKernel(nw->fn, (unsigned)grid, 1, 1, B, 1, 1, (unsigned)lds, c->stream, args, nullptr));
  return SQ_OK;
}

}  // extern "C"
