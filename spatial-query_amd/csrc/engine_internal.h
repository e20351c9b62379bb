// engine_internal.h -- objects and helpers shared by the engine's translation
// units (engine.hip: single-level trees; tlas.hip: two-level trees).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <vector>

#include "../../include/sq.h"
#include "host_tree.h"
#include "traverse.h"

using namespace sqd;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

extern thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail(SQ_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// objects
// ---------------------------------------------------------------------------

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() { if (p) (void)hipFree(p); }
  int ensure(size_t b) {
    if (b <= bytes) return SQ_OK;
    if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
    if (b == 0) return SQ_OK;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) return fail(SQ_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    bytes = b;
    return SQ_OK;
  }
  template <class T> T* as() const { return (T*)p; }
};

constexpr int SQ_BUILD_BUFS = 32;

struct sq_ctx {
  int device = 0;
  int cus = 256;                   // compute units of the device
  hipStream_t own = nullptr;       // the stream the context created
  hipStream_t stream = nullptr;    // where work is enqueued (own, or the caller's)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int block = 128;
  int rays_per_lane = 16;          // pooled kernels: chunk of 64 * rays_per_lane queries (0 = one ray per thread)
  int lds_stack = 8;               // pooled kernel: ring-stack entries kept in LDS per lane
  int ray_order = 0;               // 1: walk rays in Morton order (order.hip)
  int descend_min = 24;            // pooled walks: pause descents below this many lanes (pool.h)
  int split_depth = 0;             // pooled collect: split queries this many levels down (0 auto, -1 never)
  int wide = 2;                    // pooled nearest / any: walk the 4-wide records (traverse.h quad_records);
                                   // 2 = for the walks in the reference's order only
  int packet = 0;                  // nearest rays: 1 = one wave-coherent walk per 64 rays (packet.hip)
  int lookahead = 1;               // pooled walks in the reference's order: 1 = look-ahead form (walk_step_spec)
  int lookahead_pause = 40;        // descend_min of the look-ahead walk
  int hit_mode = SQ_HITS_REFERENCE;   // SQ_HITS_REFERENCE or SQ_HITS_GUARDED (sq.h)
  bool timed = false;
  // {next chunk, retired waves} of the pooled single-call launches: zero at
  // creation, and each launch leaves them zero again (pool_retire)
  DevBuf pool_ctr;
  int refs = 1;                    // the caller's handle + one per live tree
  DevBuf counts, scratch, keys, keys2;   // reusable temporaries of the CSR queries
  DevBuf spill;                    // ring-stack spill columns of the pooled kernel
  DevBuf okeys, okeys2, oidx, oorder, otemp;   // Morton ray order
  DevBuf counter;                  // chunk counter of the persistent pooled kernels (+ staging words)
  DevBuf stage, stage_used;        // staging blocks of the pooled collect; grow to the largest total seen
  DevBuf items, items_used;        // sub-walk blocks of the pooled collect's first expand stage, their fill
  DevBuf items2, items2_used;      // and of its second stage
  DevBuf items_cnt;                // entry count, then base, of each sub-walk the collect walks
  DevBuf build[SQ_BUILD_BUFS];     // scratch of device builds (build.hip)
};

// Drops one reference; the last one frees the stream, events and buffers.
void ctx_release(sq_ctx* c);
inline void ctx_retain(sq_ctx* c) { ++c->refs; }

struct sq_bvh {
  sq_ctx* ctx = nullptr;
  int dim = 0, kind = 0;
  uint64_t n_prims = 0;
  sqh::RefTree tree;               // host copy of the reference tree (device builds: see host_tree_ok)
  sqh::Packed packed;
  DevBuf nodes, prims, ids, root, layers, levels;
  DevBuf quads;                    // 4-wide records, rebuilt before each 4-wide walk
  sq_tree_stats stats{};
  // device builds (build.hip): the reference layout stays on the device, and
  // tree, packed.slot_of_ref / nodes and stats.quality are filled on first use
  bool host_tree_ok = true;
  uint64_t ref_nodes = 0;
  DevBuf ref_ca, ref_cb, ref_pc;
  uint32_t consumer_pf = 0;        // trees over consumer primitives (rtc.hip): floats per primitive
  double auto_baseline = 0.0;      // sq_bvh_refit_auto: surface sum of the tree as last built (0 = not yet taken)
};

// Exclusive scan of n counts into out[0..n] (out[n] = total) on the context's
// stream; temps keeps the scan's temporaries alive until the caller is done.
int scan_counts(sq_ctx* c, const uint32_t* in, uint64_t n, uint64_t* out, std::vector<std::unique_ptr<DevBuf>>& temps);
// A consumer's source compiled at run time into a loaded module (rtc.hip).
int rtc_load_module(sq_ctx* ctx, const char* header, const char* source, const char* kernels, int dim, uint32_t pf,
                    hipModule_t* out);

// TreeStats of a host reference tree (tree.rs:81-110, bvh.rs:477-489).
void tree_stats_of(const sqh::RefTree& t, sq_tree_stats& st);

// Downloads the reference tree of a device-built tree into b->tree (and the
// host arrays that depend on it); no-op for trees built on the host.
int ensure_host_tree(sq_bvh* b);

// ---------------------------------------------------------------------------
// dispatch over (D, KIND)
// ---------------------------------------------------------------------------

template <class F> inline int dispatch_dk(int dim, int kind, F&& f) {
  if (kind == KIND_TRI) {
    if (dim != 3) return fail(SQ_ERR_UNSUPPORTED, "TRIANGLE primitives need dim 3");
    return f.template operator()<3, KIND_TRI>();
  }
  if (kind != KIND_AABB && kind != KIND_SPHERE) return fail(SQ_ERR_INVALID, "unknown primitive kind");
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

// Instance kinds of a two-level tree: BOX in D = 2..6, MESH and BALL in D = 3.
template <class F> inline int dispatch_inst(int dim, int kind, F&& f) {
  if (kind == INST_MESH) {
    if (dim != 3) return fail(SQ_ERR_UNSUPPORTED, "mesh instances need dim 3");
    return f.template operator()<3, INST_MESH>();
  }
  if (kind == INST_BALL) {
    if (dim != 3) return fail(SQ_ERR_UNSUPPORTED, "ball instances need dim 3");
    return f.template operator()<3, INST_BALL>();
  }
  if (kind != INST_BOX) return fail(SQ_ERR_INVALID, "unknown instance kind");
  switch (dim) {
    case 2: return f.template operator()<2, INST_BOX>();
    case 3: return f.template operator()<3, INST_BOX>();
    case 4: return f.template operator()<4, INST_BOX>();
    case 5: return f.template operator()<5, INST_BOX>();
    case 6: return f.template operator()<6, INST_BOX>();
    default: return fail(SQ_ERR_UNSUPPORTED, "dimension not instantiated (2..6)");
  }
}

// A two-level tree's device view and sizes (tlas.hip).
struct TlasInfo {
  sq_ctx* ctx;
  int dim, kind;
  uint64_t n_inst;
  uint32_t s_top, s_bot;   // stack depths of the top tree and the deepest mesh tree
  TlasView view;
};
void tlas_info(const sq_tlas* t, TlasInfo* out);


inline uint32_t grid_for(uint64_t n, int block) { return (uint32_t)((n + block - 1) / block); }

// Rays per lane of a pooled launch over n queries: the tuned pool size, halved
// while there would be fewer chunks than waves the chip holds resident (32 per
// compute unit: 8 per SIMD), so every resident wave gets a chunk of its own.
inline uint32_t pooled_per_lane(const sq_ctx* c, uint64_t n) {
  uint32_t per = (uint32_t)(c->rays_per_lane > 0 ? c->rays_per_lane : 1);
  const uint64_t want = 32ull * (uint64_t)c->cus;
  while (per > 1 && (n + 64ull * per - 1) / (64ull * per) < want) per >>= 1;
  return per;
}

// Levels at which a pooled collect splits its queries into sub-walks.  By
// default only overlap queries split (ray and cast walks are short and even:
// splitting them costs more than it balances), and only calls of fewer than
// 2M queries (more keep the pool's lanes busy by themselves): three quarters
// of the way down the tree, where the expand stages still run balanced and
// the sub-walks left are short (measured best on 8K-262K queries of a depth-19
// tree).
inline uint32_t split_levels(const sq_ctx* c, uint64_t n, uint64_t tree_depth, bool overlap) {
  const int L = c->split_depth;
  if (L < 0 || n == 0) return 0;
  if (L > 0) return (uint32_t)L;
  if (!overlap || n >= (1ull << 21)) return 0;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((3 * tree_depth + 2) / 4, 24));
}

// Resident blocks per CU of a kernel at a block size and LDS size, queried
// once per (kernel, device, block, LDS) and then cached: the query is host
// work every pooled launch would otherwise repeat.
inline int blocks_per_cu(const void* kern, int device, int block, size_t lds, int* out) {
  struct Key {
    const void* k; int dev, block; size_t lds;
    bool operator<(const Key& o) const {
      if (k != o.k) return k < o.k;
      if (dev != o.dev) return dev < o.dev;
      if (block != o.block) return block < o.block;
      return lds < o.lds;
    }
  };
  static std::mutex mu;
  static std::map<Key, int> cache;
  const Key key{kern, device, block, lds};
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) { *out = it->second; return SQ_OK; }
  }
  int n = 0;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, block, lds));
  std::lock_guard<std::mutex> g(mu);
  cache[key] = n;
  *out = n;
  return SQ_OK;
}

inline SceneView view_of(const sq_bvh* b) {
  SceneView s;
  s.nodes = b->nodes.as<float>();
  s.prims = b->prims.as<float>();
  s.ids = b->ids.as<uint32_t>();
  s.root = b->root.as<float>();
  return s;
}

inline Filter filter_of(const sq_bvh* b, const sq_filter* f, bool* filt) {
  Filter out;
  out.layers = b->layers.as<uint64_t>();
  out.mask = f ? f->layer_mask : ~0ull;
  out.exclude = f ? f->exclude : -1;
  *filt = out.layers != nullptr || out.mask != ~0ull || out.exclude >= 0;
  return out;
}

// True while the context stream records a graph: timing events are then left
// out (the enqueued work stays the same).
inline bool capturing(const sq_ctx* c) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(c->stream, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

inline int begin_timing(sq_ctx* c) {
  if (capturing(c)) {
    c->timed = false;
    return SQ_OK;
  }
  if (!c->ev0) {
    HIP_TRY(hipEventCreate(&c->ev0));
    HIP_TRY(hipEventCreate(&c->ev1));
  }
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  return SQ_OK;
}
inline int end_timing(sq_ctx* c) {
  if (capturing(c)) return SQ_OK;
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  c->timed = true;
  return SQ_OK;
}

inline int upload(DevBuf& d, const void* h, size_t bytes, hipStream_t s) {
  int rc = d.ensure(bytes);
  if (rc) return rc;
  if (bytes) HIP_TRY(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, s));
  return SQ_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

// Morton order of n queries (origin and direction in the first 2 dim floats
// of each qs-float record; origins quantized in box = [min dim][max dim],
// device): *order (context buffer) lists query indices in key order.
int ray_order(sq_ctx* c, int dim, const float* q, uint32_t n, uint32_t qs, const uint32_t** order);

// Writes misses (leaf -1, toi +inf, zero point/normal, sub -1, any 0).
int fill_miss(sq_ctx* c, uint64_t n, int dim, NearestOut out);

// Nearest hits by the wave-coherent packet walk (packet.hip; no filter):
// exact = the reference's hits and order, else guarded hits.
int packet_nearest(sq_bvh* b, const float* q, uint64_t n, float max_toi, const float* tmax, NearestOut out, bool exact,
                   const uint32_t* order);

// Two-level tree access for the learned proposal (tlas.hip): the top tree over
// the instances' world boxes, the dimension, and the seeded walk over guarded
// hits (proposed top-tree subtrees first, then the full walk).
sq_bvh* tlas_top_tree(sq_tlas* t);
int tlas_dim(const sq_tlas* t);
int tlas_nearest_proposed(sq_tlas* t, const float* rays, uint64_t n, float max_toi, const uint16_t* top, const uint32_t* sub_ref,
                          uint32_t n_sub, int32_t* out_inst, float* out_toi, int32_t* out_sub, uint64_t* stats);

// Variable-length (CSR) results: count pass, scan, capacity check, fill pass,
// per-query sort by key, unpack.  launch(fill, keys) enqueues the count pass
// (fill = false, writing ctx->counts) or the fill pass (fill = true).
int csr_run(sq_ctx* c, uint64_t n, uint64_t* offsets, int32_t* out_leaf, float* out_toi, bool has_t,
            uint64_t capacity, uint64_t* total, const std::function<int(bool, uint64_t*)>& launch);

// Where staged entries of sub-walks find their query and place (null: the
// entries carry their query and index).
namespace sqd { struct WorkItem; }
struct StagedWalk {
  const sqd::WorkItem* items;
  const uint32_t* item_base;
};

// Variable-length results from one pooled walk per pass (pool_collect):
// launch(want_entries, want_blocks, walk) enqueues a walk appending to the
// context's staging blocks (ctx->stage, ctx->stage_used; room for at least
// want_entries), per-query counts to ctx->counts and the words of ctx->counter
// ([1] staging blocks claimed, [2] staging overflow, [4] item overflow, [6] and
// [8] item blocks claimed); then scan, capacity check, scatter, per-query sort,
// unpack, with the walk repeated when its staging or items did not fit.
int csr_run_staged(sq_ctx* c, uint64_t n, uint64_t* offsets, int32_t* out_leaf, float* out_toi, bool has_t,
                   uint64_t capacity, uint64_t* total, const std::function<int(uint64_t, uint64_t, StagedWalk&)>& launch);
