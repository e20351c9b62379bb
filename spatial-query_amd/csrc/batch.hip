// batch.hip -- many nearest-hit batches over many scenes in one persistent,
// work-stealing launch (sq_batch_*).  The reference answers one batch per call
// (raycast_nearest_batch, crates/spatial-query/src/accel/bvh.rs:134-193); here
// the batches of a frame become one table of 64 * rays_per_lane-ray chunks, and
// a grid sized to what the chip holds resident pulls chunks from an atomic
// counter until none are left, so uneven batch sizes leave no CU idle.
// Enqueueing is capture-safe (kernels only: the launch leaves its counters
// cleared for the next one), so a frame of submissions can be recorded into a
// hipGraph.
#include <algorithm>

#include "engine_internal.h"
#include "pool.h"

namespace {

struct BvhItem {
  SceneView s;
  const float* rays;
  int32_t* leaf;
  float* toi;
  uint32_t chunk0, n;     // first chunk of this batch, rays
  float max_toi;
  uint32_t pad;
};

struct TlasItem {
  TlasView tv;
  const float* rays;
  int32_t* inst;
  float* toi;
  uint32_t chunk0, n;
  float max_toi;
  uint32_t pad;
};

// Batch of chunk c: the last item whose first chunk is <= c (items hold >= 1
// chunk each).  c and the items are wave-uniform, so this runs on scalar loads.
template <class Item>
__device__ __forceinline__ uint32_t find_item(const Item* __restrict__ items, uint32_t n_items, uint32_t c) {
  uint32_t lo = 0, hi = n_items - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (items[mid].chunk0 <= c) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// EXACT: the reference's hits (pool_walk); else the guarded kinds, near-first.
template <int D, int KIND, bool EXACT>
__global__ __launch_bounds__(256, (D <= 3 ? 8 : 4)) void k_batch_bvh(const BvhItem* __restrict__ items, uint32_t n_items,
                                                                     uint32_t n_chunks, uint32_t per_lane, uint32_t S,
                                                                     SpillEntry* spill, uint32_t spill_stride, uint32_t* counter,
                                                                     uint32_t pause_below) {
  extern __shared__ uint32_t lds[];
  const uint32_t B = blockDim.x, tid = threadIdx.x;
  RingStack st{lds + tid, reinterpret_cast<float*>(lds + S * B) + tid, B, S, spill + blockIdx.x * B + tid, spill_stride};
  const Filter nof{nullptr, ~0ull, -1};
  const uint64_t pool = 64ull * per_lane;
  // a wave walks the chunks of one batch as one pool (lanes refilled across
  // its chunks) until a fetch lands in a later batch, then goes on with that
  uint32_t c = wave_fetch_add(counter);
  while (c < n_chunks) {
    const BvhItem& it = items[find_item(items, n_items, c)];
    const uint32_t end = it.chunk0 + (uint32_t)(((uint64_t)it.n + pool - 1) / pool);
    const NearestOut out{it.leaf, it.toi, nullptr, nullptr, nullptr, nullptr};
    pool_walk<D, KIND, false, false, false, ItemChunks, false, EXACT>(it.s, it.rays, ItemChunks{counter, it.chunk0, end, it.n, pool, &c}, it.max_toi,
                                            nullptr, nof, out, st, nullptr, pause_below);
  }
  pool_retire(counter);
}

// Two-level scenes: the resumable two-level walk (pool_walk_tlas) over each
// chunk, both stacks S-entry LDS rings spilling to their own columns.  EXACT as
// pool_walk_tlas.
template <int D, int IK, bool EXACT>
__global__ __launch_bounds__(256) void k_batch_tlas(const TlasItem* __restrict__ items, uint32_t n_items, uint32_t n_chunks,
                                                    uint32_t per_lane, uint32_t S, SpillEntry* spill_top, SpillEntry* spill_bot,
                                                    uint32_t spill_stride, uint32_t* counter, uint32_t pause_below) {
  extern __shared__ uint32_t lds[];
  const uint32_t B = blockDim.x, tid = threadIdx.x;
  const uint32_t gt = blockIdx.x * B + tid;
  RingStack st{lds + tid, reinterpret_cast<float*>(lds + S * B) + tid, B, S, spill_top + gt, spill_stride};
  RingStack sb{lds + 2 * S * B + tid, reinterpret_cast<float*>(lds + 3 * S * B) + tid, B, S, spill_bot + gt, spill_stride};
  const uint64_t pool = 64ull * per_lane;
  uint32_t c = wave_fetch_add(counter);
  while (c < n_chunks) {
    const TlasItem& it = items[find_item(items, n_items, c)];
    const uint32_t end = it.chunk0 + (uint32_t)(((uint64_t)it.n + pool - 1) / pool);
    const NearestOut out{it.inst, it.toi, nullptr, nullptr, nullptr, nullptr};
    pool_walk_tlas<D, IK, false, ItemChunks, EXACT>(it.tv, it.rays, ItemChunks{counter, it.chunk0, end, it.n, pool, &c}, it.max_toi, nullptr, out,
                                 st, sb, pause_below);
  }
  pool_retire(counter);
}

}  // namespace

struct sq_batch {
  sq_ctx* ctx = nullptr;
  int cls = 0, dim = 0, kind = 0;
  uint32_t n_items = 0, n_chunks = 0, per_lane = 1, S = 8, s_top = 0, s_bot = 0, grid = 0, block = 128, descend_min = 0;
  uint32_t rows_top = 0;           // two-level sets: spill rows of the top stack (the BLAS stack's follow)
  size_t lds = 0;
  DevBuf items, counter, spill;
  std::vector<std::pair<uint64_t, NearestOut>> empty;   // batches over scenes without primitives: all misses
  // BVH scenes as the table holds them: a rebuild that moved a scene's arrays
  // or deepened it past the spill rows sized here is refused at submit
  std::vector<std::pair<const sq_bvh*, SceneView>> scenes;
  uint32_t stack_cap = 0;
  bool exact = true;               // the context's hit mode when the set was made
};

// Call fb<D, K, EXACT>() (BVH sets) or ft<D, IK, EXACT>() (two-level sets) for
// the set's dimension, kind and hit mode (AABB leaves have one walk, exact in
// both modes).
template <class FB, class FT> static int dispatch_set(const sq_batch* b, FB&& fb, FT&& ft) {
  if (b->cls == SQ_BATCH_BVH)
    return dispatch_dk(b->dim, b->kind, [&]<int D, int K>() -> int {
      if constexpr (K == KIND_AABB) return fb.template operator()<D, K, true>();
      else return b->exact ? fb.template operator()<D, K, true>() : fb.template operator()<D, K, false>();
    });
  return dispatch_inst(b->dim, b->kind, [&]<int D, int K>() -> int {
    return b->exact ? ft.template operator()<D, K, true>() : ft.template operator()<D, K, false>();
  });
}


extern "C" {

int sq_batch_create(sq_ctx* ctx, int scene_class, const sq_batch_query* qs, uint32_t n_queries, sq_batch** out) {
  if (!ctx || !out || (n_queries && !qs)) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (scene_class != SQ_BATCH_BVH && scene_class != SQ_BATCH_TLAS) return fail(SQ_ERR_INVALID, "unknown scene class");
  auto b = std::make_unique<sq_batch>();
  b->ctx = ctx;
  b->cls = scene_class;
  b->per_lane = (uint32_t)std::max(1, ctx->rays_per_lane);
  b->S = (uint32_t)ctx->lds_stack;
  b->block = (uint32_t)std::min(ctx->block, 256);
  b->descend_min = (uint32_t)ctx->descend_min;
  b->exact = ctx->hit_mode != SQ_HITS_GUARDED;
  const uint64_t pool = 64ull * b->per_lane;
  std::vector<BvhItem> bi;
  std::vector<TlasItem> ti;
  uint64_t chunks = 0;
  uint32_t max_stack = 0;
  for (uint32_t i = 0; i < n_queries; ++i) {
    const sq_batch_query& q = qs[i];
    if (!q.scene) return fail(SQ_ERR_INVALID, "batch " + std::to_string(i) + ": NULL scene");
    if (q.n && (!q.rays || !q.out_leaf || !q.out_toi)) return fail(SQ_ERR_INVALID, "batch " + std::to_string(i) + ": NULL buffer");
    if (q.n > 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "batch " + std::to_string(i) + ": at most 2^32-1 rays");
    int dim, kind;
    bool empty;
    if (scene_class == SQ_BATCH_BVH) {
      const sq_bvh* s = static_cast<const sq_bvh*>(q.scene);
      if (s->ctx != ctx) return fail(SQ_ERR_INVALID, "batch " + std::to_string(i) + ": scene of another context");
      dim = s->dim; kind = s->kind; empty = s->n_prims == 0;
      if (!empty) max_stack = std::max(max_stack, s->packed.max_stack);
    } else {
      const sq_tlas* t = static_cast<const sq_tlas*>(q.scene);
      TlasInfo inf;
      tlas_info(t, &inf);
      if (inf.ctx != ctx) return fail(SQ_ERR_INVALID, "batch " + std::to_string(i) + ": scene of another context");
      dim = inf.dim; kind = inf.kind; empty = inf.n_inst == 0;
      if (!empty) { b->s_top = std::max(b->s_top, inf.s_top); b->s_bot = std::max(b->s_bot, inf.s_bot); }
    }
    if (i == 0) { b->dim = dim; b->kind = kind; }
    if (dim != b->dim || kind != b->kind)
      return fail(SQ_ERR_INVALID, "batch " + std::to_string(i) + ": every scene of a batch set needs the same dimension and kind");
    if (!q.n) continue;
    if (empty) {
      b->empty.push_back({q.n, NearestOut{q.out_leaf, q.out_toi, nullptr, nullptr, nullptr, nullptr}});
      continue;
    }
    const uint32_t c0 = (uint32_t)chunks;
    chunks += (q.n + pool - 1) / pool;
    if (chunks > 0x7FFFFFFFull) return fail(SQ_ERR_INVALID, "too many chunks in one batch set");
    if (scene_class == SQ_BATCH_BVH) {
      BvhItem it{};
      it.s = view_of(static_cast<const sq_bvh*>(q.scene));
      b->scenes.emplace_back(static_cast<const sq_bvh*>(q.scene), it.s);
      it.rays = q.rays; it.leaf = q.out_leaf; it.toi = q.out_toi;
      it.chunk0 = c0; it.n = (uint32_t)q.n; it.max_toi = q.max_toi;
      bi.push_back(it);
    } else {
      TlasInfo inf;
      tlas_info(static_cast<const sq_tlas*>(q.scene), &inf);
      TlasItem it{};
      it.tv = inf.view;
      it.rays = q.rays; it.inst = q.out_leaf; it.toi = q.out_toi;
      it.chunk0 = c0; it.n = (uint32_t)q.n; it.max_toi = q.max_toi;
      ti.push_back(it);
    }
  }
  b->n_items = (uint32_t)(scene_class == SQ_BATCH_BVH ? bi.size() : ti.size());
  b->n_chunks = (uint32_t)chunks;
  int rc;
  if ((rc = b->counter.ensure(256))) return rc;
  if (b->n_items) {
    // grid: every wave the chip holds resident, capped by the chunk count
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    int per_cu = 0;
    b->lds = (size_t)(scene_class == SQ_BATCH_BVH ? 1 : 2) * b->S * b->block * 8;
    rc = dispatch_set(
        b.get(),
        [&]<int D, int K, bool EXACT>() -> int {
          HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_batch_bvh<D, K, EXACT>),
                                                               (int)b->block, b->lds));
          return SQ_OK;
        },
        [&]<int D, int K, bool EXACT>() -> int {
          HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_batch_tlas<D, K, EXACT>),
                                                               (int)b->block, b->lds));
          return SQ_OK;
        });
    if (rc) return rc;
    const uint64_t waves_per_block = b->block / 64;
    const uint64_t need = (chunks + waves_per_block - 1) / waves_per_block;
    b->grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)cus * std::max(1, per_cu), need));
    if (scene_class == SQ_BATCH_BVH) {
      b->stack_cap = max_stack;
      const uint32_t rows = max_stack > b->S ? max_stack - b->S : 0;
      if ((rc = b->spill.ensure((size_t)rows * b->grid * b->block * sizeof(SpillEntry) + 64))) return rc;
      if ((rc = upload(b->items, bi.data(), bi.size() * sizeof(BvhItem), ctx->stream))) return rc;
    } else {
      // spill columns: the top stack's rows, then the BLAS stack's
      const uint64_t lanes = (uint64_t)b->grid * b->block;
      b->rows_top = b->s_top > b->S ? b->s_top - b->S : 0;
      const uint32_t rows_bot = b->s_bot > b->S ? b->s_bot - b->S : 0;
      if ((rc = b->spill.ensure((size_t)(b->rows_top + rows_bot) * lanes * sizeof(SpillEntry) + 64))) return rc;
      if ((rc = upload(b->items, ti.data(), ti.size() * sizeof(TlasItem), ctx->stream))) return rc;
    }
    // the pool counters start zero; every launch leaves them zero (pool_retire)
    HIP_TRY(hipMemsetAsync(b->counter.p, 0, 256, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));   // the host tables end with this call
  }
  ctx_retain(ctx);
  *out = b.release();
  return SQ_OK;
}

int sq_batch_check(sq_batch* b) {
  if (!b) return fail(SQ_ERR_INVALID, "NULL argument");
  for (const auto& sv : b->scenes) {
    const SceneView now = view_of(sv.first);
    if (now.nodes != sv.second.nodes || now.prims != sv.second.prims || now.ids != sv.second.ids ||
        now.root != sv.second.root || sv.first->packed.max_stack > b->stack_cap)
      return fail(SQ_ERR_INVALID, "a scene of this batch set was rebuilt into new arrays or deeper; make the batch set again");
  }
  return SQ_OK;
}

int sq_batch_submit(sq_batch* b) {
  int rc = sq_batch_check(b);
  if (rc) return rc;
  sq_ctx* c = b->ctx;
  rc = begin_timing(c);
  if (rc) return rc;
  for (const auto& e : b->empty)
    if ((rc = fill_miss(c, e.first, b->dim, e.second))) return rc;
  if (b->n_items) {
    const uint32_t spill_stride = b->grid * b->block;
    SpillEntry* sp = b->spill.as<SpillEntry>();
    rc = dispatch_set(
        b,
        [&]<int D, int K, bool EXACT>() -> int {
          k_batch_bvh<D, K, EXACT><<<b->grid, b->block, b->lds, c->stream>>>(b->items.as<BvhItem>(), b->n_items, b->n_chunks,
                                                                            b->per_lane, b->S, sp, spill_stride,
                                                                            b->counter.as<uint32_t>(), b->descend_min);
          HIP_TRY(hipGetLastError());
          return SQ_OK;
        },
        [&]<int D, int K, bool EXACT>() -> int {
          k_batch_tlas<D, K, EXACT><<<b->grid, b->block, b->lds, c->stream>>>(b->items.as<TlasItem>(), b->n_items, b->n_chunks,
                                                                             b->per_lane, b->S, sp,
                                                                             sp + (size_t)b->rows_top * spill_stride, spill_stride,
                                                                             b->counter.as<uint32_t>(), b->descend_min);
          HIP_TRY(hipGetLastError());
          return SQ_OK;
        });
    if (rc) return rc;
  }
  return end_timing(c);
}

int sq_batch_get_info(const sq_batch* b, uint32_t* n_chunks, uint32_t* grid, uint32_t* block) {
  if (!b) return fail(SQ_ERR_INVALID, "NULL argument");
  if (n_chunks) *n_chunks = b->n_chunks;
  if (grid) *grid = b->grid;
  if (block) *block = b->block;
  return SQ_OK;
}

int sq_batch_destroy(sq_batch* b) {
  if (!b) return SQ_OK;
  sq_ctx* c = b->ctx;
  (void)hipStreamSynchronize(c->stream);
  delete b;
  ctx_release(c);
  return SQ_OK;
}

}  // extern "C"
