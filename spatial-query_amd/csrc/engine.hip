// engine.hip -- HIP kernels (gfx950) and the C ABI of include/sq.h.
//
// One lane per query, 64-lane wavefronts, a per-lane traversal stack in LDS
// laid out entry-major (entry k of lane l at k * blockDim + l) so a wave's
// pushes and pops hit 64 consecutive banks.  The per-ray logic lives in
// traverse.h; this file owns device memory, launches and the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine_internal.h"
#include "pool.h"

thread_local std::string g_err;

// Device memory of sq_raycast_nearest_noalloc: the pooled walk's two counters
// (the first 256 bytes) and its ring-stack spill columns (the rest), in place
// of the context's buffers.  Each device holds its own copy until the process
// ends; its address there comes from hipGetSymbolAddress (g_scratch below),
// since host code cannot use the array's name as a device address.
static __device__ __align__(256) unsigned char scratch[SQ_NOALLOC_SCRATCH_BYTES];
constexpr int SQ_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch[SQ_MAX_DEVICES];   // scratch's address on each device, set once by scratch_init

// Looks up scratch's address on `device`, the current device (once per device).
static int scratch_init(int device) {
  if (device < 0 || device >= SQ_MAX_DEVICES) return fail(SQ_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch[device]) return SQ_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(SQ_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[device] = p;
  return SQ_OK;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Nearest / any-hit rays (RayPolicy) and nearest shape casts (CastPolicy).
// Query i reads QS floats at q + i*QS.  EXACT: the reference's hits (unguarded
// narrow phases, nearest queries whose hits may leave their box in the
// reference's order); else the guarded kinds, near-first.
template <int D, int KIND, bool ANY, bool FILT, bool CAST, bool EXACT>
__global__ __launch_bounds__(1024) void k_nearest(SceneView s, const float* __restrict__ q, uint32_t n, float max_toi,
                                                  const float* __restrict__ tmax, Filter f, NearestOut out, uint32_t stack_n) {
  extern __shared__ uint32_t lds[];
  const uint32_t B = blockDim.x, tid = threadIdx.x;
  const uint32_t i = blockIdx.x * B + tid;
  if (i >= n) return;
  using Pol = typename std::conditional<CAST, CastPolicy<D, KIND, float, !EXACT>, RayPolicy<D, KIND, !EXACT>>::type;
  constexpr int QS = CAST ? 4 * D : 2 * D;
  constexpr int PS = Prim<D, KIND>::STRIDE;
  Pol pol;
  pol.setup(q + (size_t)i * QS);
  const float lim = tmax ? tmax[i] : max_toi;
  Stack st{lds + tid, reinterpret_cast<float*>(lds + stack_n * B) + tid, B};
  NearestResult h = walk_nearest<D, PS, ANY, FILT, EXACT && !ANY && !Pol::INSIDE>(s, pol, lim, f, st);
  if (ANY) { out.any[i] = h.leaf != NO_LEAF ? 1 : 0; return; }
  const bool hit = h.leaf != NO_LEAF;
  out.leaf[i] = hit ? (int32_t)h.leaf : -1;
  out.toi[i] = h.t;
  if (out.point) {
    // Ray::at / ShapeCast::at: origin + dir * t
#pragma unroll
    for (int k = 0; k < D; ++k) out.point[(size_t)i * D + k] = hit ? pol.r.o[k] + pol.r.d[k] * h.t : 0.0f;
  }
  if (!CAST && out.normal) {
    float nv[D];
    if (hit) hit_normal<D, KIND>(pol.r, s.prims + (size_t)h.pos * PS, h.t, lim, nv);
    else for (int k = 0; k < D; ++k) nv[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < D; ++k) out.normal[(size_t)i * D + k] = nv[k];
  }
  if (out.sub) out.sub[i] = (hit && KIND != KIND_AABB) ? (int32_t)h.leaf : -1;
}


__global__ void k_zero_u32(uint32_t* p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// The reference's QueryStats per ray (walk_counted): nearest (ALL = false,
// leaf and toi written when out_leaf is given) or all hits (ALL).  One ray per
// thread, grid-stride; each thread's stack is a column of `stack` (refs, then
// entries), `lanes` threads wide.
template <int D, int KIND, bool ALL, bool FILT>
__global__ __launch_bounds__(256) void k_counted(SceneView s, const float* __restrict__ q, uint32_t n, float max_toi,
                                                 const float* __restrict__ tmax, Filter f, int32_t* out_leaf, float* out_toi,
                                                 uint32_t* out_stats, uint32_t* stack, uint32_t stack_n) {
  const uint32_t lanes = gridDim.x * blockDim.x;
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  constexpr int PS = Prim<D, KIND>::STRIDE;
  Stack st{stack + tid, reinterpret_cast<float*>(stack + (size_t)stack_n * lanes) + tid, lanes};
  for (uint32_t i = tid; i < n; i += lanes) {
    RayPolicy<D, KIND, false> pol;
    pol.setup(q + (size_t)i * 2 * D);
    uint32_t cnt[4];
    const NearestResult h = walk_counted<D, PS, ALL, FILT>(s, pol, tmax ? tmax[i] : max_toi, f, st, cnt);
#pragma unroll
    for (int k = 0; k < 4; ++k) out_stats[(size_t)i * 4 + k] = cnt[k];
    if (!ALL && out_leaf) {
      out_leaf[i] = h.leaf != NO_LEAF ? (int32_t)h.leaf : -1;
      out_toi[i] = h.t;
    }
  }
}

// Pooled nearest / any-hit / shape-cast kernel, persistent: the grid holds as
// many waves as the chip keeps resident, and every wave takes chunks of `chunk`
// queries from counter[0] until all n are handed out, refilling its lanes
// across chunk boundaries (pool_walk); the counters start zero and the launch
// leaves them zero (pool_retire).  With `order`, pool
// position i is query order[i] (the Morton order).  QUAD walks the 4-wide
// records s.quads (k_quads) instead of the pair nodes.  EXACT as k_nearest.
// SPEC: the reference-order walks take the look-ahead form (walk_step_spec),
// held to 7 waves per SIMD (its pending leaf and position need registers the
// 8-wave budget of 64 would spill: C1 5.68 against 5.81 ms; 6 waves 6.41).
template <int D, int KIND, bool ANY, bool FILT, bool CAST, bool QUAD, bool EXACT, bool SPEC = false>
__global__ __launch_bounds__(256, (SPEC ? 7 : (D <= 3 && !CAST && !FILT ? 8 : 4))) void k_nearest_pool(SceneView s, const float* __restrict__ q, uint32_t n, float max_toi,
                                                      const float* __restrict__ tmax, Filter f, NearestOut out,
                                                      uint32_t chunk, uint32_t* counter, uint32_t S, SpillEntry* spill,
                                                      uint32_t spill_stride, const uint32_t* __restrict__ order,
                                                      uint32_t pause_below) {
  extern __shared__ uint32_t lds[];
  const uint32_t B = blockDim.x, tid = threadIdx.x;
  const uint32_t gtid = blockIdx.x * B + tid;
  RingStack st{lds + tid, reinterpret_cast<float*>(lds + S * B) + tid, B, S, spill + gtid, spill_stride};
  const uint64_t waves = (uint64_t)gridDim.x * (B >> 6);
  const uint64_t home = (uint64_t)blockIdx.x * (B >> 6) + (tid >> 6);
  pool_walk<D, KIND, ANY, FILT, CAST, Chunks, QUAD, EXACT, SPEC>(s, q, Chunks{counter, n, chunk, home, waves}, max_toi, tmax,
                                                               f, out, st, order, pause_below);
  pool_retire(counter);
}

// The 4-wide records of every pair node (quad_records), rebuilt from the pair
// nodes before each QUAD walk so they follow refits and rebuilds.
template <int D>
__global__ void k_quads(const float* __restrict__ nodes, uint32_t n_nodes, float* __restrict__ quads) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_nodes) return;
  constexpr int NF = (D + 1) * 4;
  quad_records<D>(nodes, n_nodes, p, quads + (size_t)p * 2 * NF);
}

template <int D>
__global__ __launch_bounds__(1024) void k_required(SceneView s, const float* __restrict__ rays, uint32_t n, const float* __restrict__ tlim,
                                                   uint32_t* out_nodes, uint32_t* out_prims, uint32_t stack_n) {
  extern __shared__ uint32_t lds[];
  const uint32_t B = blockDim.x, tid = threadIdx.x;
  const uint32_t i = blockIdx.x * B + tid;
  if (i >= n) return;
  RayD<D> r;
  ray_setup<D>(r, rays + (size_t)i * 2 * D);
  Stack st{lds + tid, nullptr, B};
  uint32_t nn, np;
  walk_required<D>(s, r, tlim[i], st, nn, np);
  out_nodes[i] = nn;
  out_prims[i] = np;
}

// --- variable-length queries: count pass, then fill pass -------------------
// MODE 0: raycast_all, 1: raycast_crossings, 2: overlap, 3: shapecast_all.
// FILL = false counts hits per query into counts[i]; FILL = true writes sort
// keys at keys[offsets[i] + j] (hit_key for rays and casts, the leaf for overlap).
template <int D, int KIND, int MODE, bool FILL, bool FILT>
__global__ __launch_bounds__(1024) void k_collect(SceneView s, const float* __restrict__ q, uint32_t n, float max_toi,
                                                  const float* __restrict__ tmax, Filter f, uint32_t* counts,
                                                  const uint64_t* __restrict__ offsets, uint64_t* keys) {
  extern __shared__ uint32_t lds[];
  const uint32_t B = blockDim.x, tid = threadIdx.x;
  const uint32_t i = blockIdx.x * B + tid;
  if (i >= n) return;
  constexpr int PS = Prim<D, KIND>::STRIDE;
  Stack st{lds + tid, nullptr, B};
  uint32_t cnt = 0;
  uint64_t* dst = FILL ? keys + offsets[i] : nullptr;
  auto emit = [&](uint64_t key) {
    if (FILL) dst[cnt] = key;
    ++cnt;
  };
  const uint32_t root_ref = fbits(s.root[2 * D]);
  if constexpr (MODE == 2) {
    const float* rg = q + (size_t)i * 2 * D;
    float rmn[D], rmx[D];
#pragma unroll
    for (int k = 0; k < D; ++k) { rmn[k] = rg[k]; rmx[k] = rg[D + k]; }
    if (box_overlap<D>(s.root, s.root + D, rmn, rmx)) {
      walk_visit<D>(s, root_ref,
        [&](const float* nd, bool& ha, bool& hb) {
          if (KIND != KIND_AABB && synthetic_pair<D>(nd)) { ha = hb = true; return; }   // parts of one reference leaf
          float amn[D], amx[D], bmn[D], bmx[D];
#pragma unroll
          for (int k = 0; k < D; ++k) { amn[k] = nd[4 * k]; amx[k] = nd[4 * k + 1]; bmn[k] = nd[4 * k + 2]; bmx[k] = nd[4 * k + 3]; }
          ha = box_overlap<D>(amn, amx, rmn, rmx);
          hb = box_overlap<D>(bmn, bmx, rmn, rmx);
        },
        [&](uint32_t first, uint32_t c) {
          for (uint32_t j = 0; j < c; ++j) {
            const uint32_t pos = first + j;
            if (!overlap_narrow<D, KIND>(s.prims + (size_t)pos * PS, rmn, rmx)) continue;
            const uint32_t leaf = s.ids[pos];
            if (FILT && !accepts(f, leaf)) continue;
            emit(leaf);
          }
        }, st);
    }
  } else {
    using Pol = typename std::conditional<MODE == 3, CastPolicy<D, KIND>, RayPolicy<D, KIND>>::type;
    constexpr int QS = MODE == 3 ? 4 * D : 2 * D;
    Pol pol;
    pol.setup(q + (size_t)i * QS);
    const float lim = tmax ? tmax[i] : max_toi;
    if (pol.root(s.root, lim)) {
      walk_visit<D>(s, root_ref,
        [&](const float* nd, bool& ha, bool& hb) {
          if (KIND != KIND_AABB && synthetic_pair<D>(nd)) { ha = hb = true; return; }   // parts of one reference leaf
          float ta, tb;
          pol.pair(nd, lim, ha, hb, ta, tb);
        },
        [&](uint32_t first, uint32_t c) {
          for (uint32_t j = 0; j < c; ++j) {
            const uint32_t pos = first + j;
            const float* p = s.prims + (size_t)pos * PS;
            float t;
            if constexpr (MODE == 1 && KIND == KIND_AABB) {
              // Box crossings: entry, and the real exit when it lies within max_toi.
              float te, tx; int ax;
              if (!slab_exact<D>(pol.r, p, p + D, lim, te, tx, ax)) continue;
              const uint32_t leaf = s.ids[pos];
              if (FILT && !accepts(f, leaf)) continue;
              emit(hit_key(te, leaf));
              if (tx > te) {
                float te2, tx2;
                slab_exact<D>(pol.r, p, p + D, SQ_INF, te2, tx2, ax);
                if (tx2 <= lim) emit(hit_key(tx2, leaf));
              }
            } else {
              if (!pol.narrow(p, lim, t)) continue;
              const uint32_t leaf = s.ids[pos];
              if (FILT && !accepts(f, leaf)) continue;
              emit(hit_key(t, leaf));
            }
          }
        }, st);
    }
  }
  if (!FILL) counts[i] = cnt;
}

// Pooled, persistent collect (pool_collect): waves take chunks of `chunk`
// queries from counter[0], or with ITEMS the filled parts of the first
// min(*item_blocks, max_item_blocks) item blocks of k_collect_expand (each
// sub-walk's entry count to item_cnt); entries go to the staging blocks of
// `o` (counter[1] = blocks claimed, counter[2] = overflow flag).  LDS: the S-entry ring stacks, then two words per wave for
// its open staging block.
template <int D, int KIND, int MODE, bool FILT, bool ITEMS>
__global__ __launch_bounds__(256) void k_collect_pool(SceneView s, const float* __restrict__ q, uint32_t n, float max_toi,
                                                      const float* __restrict__ tmax, Filter f, uint32_t chunk,
                                                      uint32_t* counter, uint32_t S, SpillEntry* spill,
                                                      uint32_t spill_stride, uint32_t pause_below, StageOut o,
                                                      const WorkItem* __restrict__ items, const uint32_t* __restrict__ item_used,
                                                      const uint32_t* item_blocks, uint32_t max_item_blocks, uint32_t* item_cnt) {
  extern __shared__ uint32_t lds[];
  const uint32_t B = blockDim.x, tid = threadIdx.x;
  const uint32_t gtid = blockIdx.x * B + tid;
  RingStack st{lds + tid, reinterpret_cast<float*>(lds + S * B) + tid, B, S, spill + gtid, spill_stride};
  volatile uint32_t* slot = lds + 2 * S * B + 2 * (tid >> 6);
  if ((tid & 63u) == 0) {
    slot[0] = NO_BLOCK;
    slot[1] = 0;
  }
  const uint64_t waves = (uint64_t)gridDim.x * (B >> 6);
  const uint64_t home = (uint64_t)blockIdx.x * (B >> 6) + (tid >> 6);
  if constexpr (ITEMS) {
    const uint64_t nb = std::min<uint32_t>(*item_blocks, max_item_blocks);
    pool_collect<D, KIND, MODE, FILT, true>(s, q, BlockChunks{counter, nb, item_used, home, waves}, max_toi, tmax, f, st,
                                            StageAppend{o, slot}, pause_below, items, item_cnt);
  } else {
    pool_collect<D, KIND, MODE, FILT, false>(s, q, Chunks{counter, n, chunk, home, waves}, max_toi, tmax, f, st,
                                             StageAppend{o, slot}, pause_below, items, item_cnt);
  }
}

// Splits queries into sub-walks: a depth-first walk that stops `levels` pair
// nodes down and appends (query, node) for every node it reaches there, and
// (query, leaf ref) for every leaf range it reaches above that, to item
// blocks.  Without FROM_ITEMS the walks start at the root of each of the n
// queries (which also zeroes counts[i], the item walks add to it); with
// FROM_ITEMS they start at the sub-walks of the first min(*in_blocks,
// in_max_blocks) blocks of in_items (a grid-stride loop over their slots).
// Each wave fills one block of ITEM_BLK at a time, claimed from *out_blocks;
// out_used[b] = the items block b holds.  Blocks at or beyond out_max_blocks
// are not written and set *overflow.  Every node and leaf range the query's
// full walk would visit lies below exactly one item, so the items' walks
// together visit the same leaf ranges.  At most 2^levels items per input.
// LDS: `levels` stack entries (ref, depth) per thread.
template <int D, int KIND, int MODE, bool FILT, bool FROM_ITEMS>
__global__ __launch_bounds__(256) void k_collect_expand(SceneView s, const float* __restrict__ q, uint32_t n, float max_toi,
                                                        const float* __restrict__ tmax, uint32_t levels,
                                                        const WorkItem* __restrict__ in_items, const uint32_t* __restrict__ in_used,
                                                        const uint32_t* in_blocks, uint32_t in_max_blocks,
                                                        WorkItem* items, uint32_t* out_used, uint32_t* out_blocks,
                                                        uint32_t out_max_blocks, uint32_t* overflow, uint32_t* counts) {
  extern __shared__ uint32_t lds[];
  const uint32_t B = blockDim.x, tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  uint32_t* sref = lds + tid;
  uint32_t* sdep = lds + levels * B + tid;
  constexpr int NF = (D + 1) * 4;
  const uint64_t n_in = FROM_ITEMS ? (uint64_t)std::min<uint32_t>(*in_blocks, in_max_blocks) * ITEM_BLK : (uint64_t)n;
  // the wave's open item block and its fill (the same in every lane: only
  // changed where the whole wave runs)
  uint32_t blk = NO_BLOCK, fill = 0;
  for (uint64_t first = (uint64_t)blockIdx.x * B; first < n_in; first += (uint64_t)gridDim.x * B) {
    const uint64_t i = first + tid;
    CollectQuery<D, KIND, MODE, FILT> cq;
    bool live = i < n_in;
    uint32_t qi = 0, cur = fbits(s.root[2 * D]);
    if constexpr (FROM_ITEMS) {
      if (live) live = (uint32_t)(i % ITEM_BLK) < in_used[i / ITEM_BLK];
      if (live) {
        const WorkItem it = in_items[i];
        qi = it.q;
        cur = it.ref;
        (void)cq.begin(s, q, qi, max_toi, tmax);        // the first expand ran the root test
      }
    } else {
      if (live) {
        qi = (uint32_t)i;
        counts[qi] = 0;
        live = cq.begin(s, q, qi, max_toi, tmax);
      }
    }
    uint32_t depth = 0, sp = 0;
    while (__ballot(live)) {
      const bool at = live && ((cur & LEAF_BIT) || depth == levels);
      const uint64_t m = __ballot(at);
      if (m) {
        const uint32_t k = (uint32_t)__popcll(m);
        if (blk == NO_BLOCK || fill + k > ITEM_BLK) {
          if (blk != NO_BLOCK && blk < out_max_blocks && lane == 0) out_used[blk] = fill;
          uint32_t nb = 0;
          if (lane == 0) nb = atomicAdd(out_blocks, 1u);
          blk = (uint32_t)__builtin_amdgcn_readfirstlane((int)nb);
          fill = 0;
          if (blk >= out_max_blocks && lane == 0) *overflow = 1u;
        }
        if (at && blk < out_max_blocks) items[(size_t)blk * ITEM_BLK + fill + (uint32_t)__popcll(m & lt_mask)] = WorkItem{qi, cur};
        fill += k;
      }
      bool pop = at;
      if (live && !at) {
        const float* nd = s.nodes + (size_t)cur * NF;
        bool ha, hb;
        cq.pair(nd, ha, hb);
        const uint32_t ra = fbits(nd[4 * D + 0]), rb = fbits(nd[4 * D + 1]);
        if (ha && hb) {
          sref[sp * B] = rb;
          sdep[sp * B] = depth + 1;
          ++sp;
          cur = ra;
          ++depth;
        } else if (ha || hb) {
          cur = ha ? ra : rb;
          ++depth;
        } else {
          pop = true;
        }
      }
      if (pop) {
        if (sp == 0) {
          live = false;
        } else {
          --sp;
          cur = sref[sp * B];
          depth = sdep[sp * B];
        }
      }
    }
  }
  if (blk != NO_BLOCK && blk < out_max_blocks && lane == 0) out_used[blk] = fill;
}

__global__ void k_zero_words(uint32_t* p, uint32_t n) {
  if (threadIdx.x < n) p[threadIdx.x] = 0u;
}

// Each sub-walk's place in its query's list: item_cnt[p] (its entry count) is
// replaced by the count of the query's entries placed before it, taken from
// counts[q] (zeroed by the expand), which ends as the query's total.  Runs
// over the filled slots of the first min(*blocks, max_blocks) item blocks.
__global__ __launch_bounds__(256) void k_item_base(const WorkItem* __restrict__ items, const uint32_t* __restrict__ used,
                                                   const uint32_t* blocks, uint32_t max_blocks, uint32_t* item_cnt,
                                                   uint32_t* counts) {
  const uint64_t slots = (uint64_t)std::min<uint32_t>(*blocks, max_blocks) * ITEM_BLK;
  for (uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x; p < slots; p += (uint64_t)gridDim.x * 256) {
    if ((uint32_t)(p % ITEM_BLK) >= used[p / ITEM_BLK]) continue;
    const uint32_t c = item_cnt[p];
    item_cnt[p] = c ? atomicAdd(counts + items[p].q, c) : 0u;
  }
}

// Staged entries to their CSR slots, for the first used[b] entries of each
// block b < n_blocks: keys[offsets[q] + idx] = key; with items, an entry's q
// is its sub-walk p and its slot is offsets[items[p].q] + item_base[p] + idx.
__global__ __launch_bounds__(256) void k_stage_scatter(const StageEntry* __restrict__ e, const uint32_t* __restrict__ used,
                                                       uint64_t n_blocks, const uint64_t* __restrict__ offsets,
                                                       const WorkItem* __restrict__ items, const uint32_t* __restrict__ item_base,
                                                       uint64_t* keys) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n_blocks * STAGE_BLK) return;
  if ((uint32_t)(i % STAGE_BLK) >= used[i / STAGE_BLK]) return;
  const StageEntry x = e[i];
  if (items) keys[offsets[items[x.q].q] + item_base[x.q] + x.idx] = x.key;
  else keys[offsets[x.q] + x.idx] = x.key;
}

// Exclusive scan of n values (u32 or u64) into out[0..n] (out[n] = total).
// Tiles of 4096 = 256 threads x 16; tile sums scanned recursively by the host.
template <class T>
__global__ __launch_bounds__(256) void k_scan_tile(const T* __restrict__ in, uint64_t n, uint64_t* out, uint64_t* tile_sums) {
  __shared__ uint64_t sh[256];
  const uint64_t base = (uint64_t)blockIdx.x * 4096 + threadIdx.x * 16;
  uint64_t v[16], run = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) { v[j] = (base + j < n) ? (uint64_t)in[base + j] : 0; run += v[j]; }
  sh[threadIdx.x] = run;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    uint64_t add = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += add;
    __syncthreads();
  }
  uint64_t acc = sh[threadIdx.x] - run;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    if (base + j <= n) out[base + j] = acc;
    acc += v[j];
  }
  if (threadIdx.x == 255) tile_sums[blockIdx.x] = sh[255];
}

__global__ void k_scan_add(uint64_t* out, uint64_t n1, const uint64_t* __restrict__ tile_off) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n1) out[i] += tile_off[i / 4096];
}

// Segment sort, small segments: one lane per segment, insertion sort in place.
// Segments longer than SMALL_SEG are appended to `large` for the block sort.
constexpr uint32_t SMALL_SEG = 32;
__global__ void k_sort_small(uint64_t* keys, const uint64_t* __restrict__ offsets, uint32_t n, uint32_t* large, uint32_t* n_large) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t a = offsets[i], b = offsets[i + 1];
  const uint64_t len = b - a;
  if (len <= 1) return;
  if (len > SMALL_SEG) { large[atomicAdd(n_large, 1u)] = i; return; }
  uint64_t* k = keys + a;
  for (uint32_t x = 1; x < len; ++x) {
    uint64_t v = k[x];
    int y = (int)x - 1;
    while (y >= 0 && k[y] > v) { k[y + 1] = k[y]; --y; }
    k[y + 1] = v;
  }
}

// Segment sort, large segments: one block per segment.  Chunks of up to 2048
// keys are bitonic-sorted in LDS (a network of the chunk's size rounded up to
// a power of two), then runs are merged pairwise (each element's output
// position = its rank in its own run + its rank in the other run),
// ping-ponging between the segment and a scratch area of the same size.
constexpr uint32_t CHUNK = 2048;
__global__ __launch_bounds__(256) void k_sort_large(uint64_t* keys, uint64_t* scratch, const uint64_t* __restrict__ offsets,
                                                    const uint32_t* __restrict__ large) {
  __shared__ uint64_t sh[CHUNK];
  const uint32_t seg = large[blockIdx.x];
  const uint64_t a = offsets[seg], len = offsets[seg + 1] - a;
  uint64_t* src = keys + a;
  uint64_t* dst = scratch + a;
  for (uint64_t c0 = 0; c0 < len; c0 += CHUNK) {
    const uint32_t m = (uint32_t)min((uint64_t)CHUNK, len - c0);
    uint32_t P = 64;
    while (P < m) P <<= 1;
    for (uint32_t j = threadIdx.x; j < P; j += blockDim.x) sh[j] = j < m ? src[c0 + j] : ~0ull;
    __syncthreads();
    for (uint32_t size = 2; size <= P; size <<= 1) {
      for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
        for (uint32_t j = threadIdx.x; j < P / 2; j += blockDim.x) {
          const uint32_t lo = 2 * j - (j & (stride - 1));
          const uint32_t hi = lo + stride;
          const bool up = ((lo & size) == 0);
          uint64_t x = sh[lo], y = sh[hi];
          if ((x > y) == up) { sh[lo] = y; sh[hi] = x; }
        }
        __syncthreads();
      }
    }
    for (uint32_t j = threadIdx.x; j < m; j += blockDim.x) src[c0 + j] = sh[j];
    __syncthreads();
  }
  // merge passes
  for (uint64_t run = CHUNK; run < len; run <<= 1) {
    for (uint64_t j = threadIdx.x; j < len; j += blockDim.x) {
      const uint64_t r0 = (j / (2 * run)) * (2 * run);
      const bool left = (j - r0) < run;
      const uint64_t mine0 = left ? r0 : r0 + run;
      const uint64_t oth0 = left ? r0 + run : r0;
      const uint64_t oth_len = left ? (min(len, r0 + 2 * run) > r0 + run ? min(len, r0 + 2 * run) - (r0 + run) : 0) : run;
      const uint64_t v = src[j];
      // rank in the other run: left elements go before equal right elements
      uint64_t lo = 0, hi = oth_len;
      while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        const uint64_t w = src[oth0 + mid];
        if (left ? (w < v) : (w <= v)) lo = mid + 1; else hi = mid;
      }
      dst[r0 + (j - mine0) + lo] = v;
    }
    __syncthreads();
    uint64_t* t = src; src = dst; dst = t;
    __syncthreads();
  }
  if (src != keys + a) {
    for (uint64_t j = threadIdx.x; j < len; j += blockDim.x) keys[a + j] = src[j];
  }
}

// Sorted keys -> caller outputs.
__global__ void k_unpack(const uint64_t* __restrict__ keys, uint64_t total, int32_t* leaf, float* toi, int has_t) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const uint64_t k = keys[i];
  leaf[i] = (int32_t)(uint32_t)k;
  if (has_t) toi[i] = key_toi(k);
}

__global__ void k_fill_miss(uint32_t n, int dim, NearestOut out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (out.any) { out.any[i] = 0; return; }
  out.leaf[i] = -1;
  out.toi[i] = SQ_INF;
  for (int k = 0; k < dim; ++k) {
    if (out.point) out.point[(size_t)i * dim + k] = 0.0f;
    if (out.normal) out.normal[(size_t)i * dim + k] = 0.0f;
  }
  if (out.sub) out.sub[i] = -1;
}

// --- refit (tree.rs:54-75 on the device) -----------------------------------

// Caller-layout primitives (leaf-indexed) -> leaf-ordered device records.
template <int D, int KIND>
__global__ void k_scatter_prims(const float* __restrict__ src, const uint32_t* __restrict__ ids, float* __restrict__ dst, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t id = ids[k];
  if constexpr (KIND == KIND_TRI) {
    const float* p = src + (size_t)id * 9;
    float* q = dst + (size_t)k * 12;
    q[0] = p[0]; q[1] = p[1]; q[2] = p[2]; q[3] = __uint_as_float(id);
    q[4] = p[3]; q[5] = p[4]; q[6] = p[5]; q[7] = 0.0f;
    q[8] = p[6]; q[9] = p[7]; q[10] = p[8]; q[11] = 0.0f;
  } else {
    constexpr int S = Prim<D, KIND>::STRIDE;
    const float* p = src + (size_t)id * S;
    float* q = dst + (size_t)k * S;
#pragma unroll
    for (int j = 0; j < S; ++j) q[j] = p[j];
  }
}

// Box of a child ref: a leaf folds its primitives' boxes from EMPTY in
// prim_indices order; a pair node is the union of its two stored boxes.
template <int D, int KIND>
__device__ __forceinline__ void child_box(const float* nodes, const float* prims, uint32_t ref, float* mn, float* mx) {
  constexpr int NF = (D + 1) * 4;
#pragma unroll
  for (int k = 0; k < D; ++k) { mn[k] = SQ_INF; mx[k] = -SQ_INF; }
  if (ref & LEAF_BIT) {
    const uint32_t first = ref & FIRST_MASK, cnt = ((ref >> 27) & 15u) + 1u;
    for (uint32_t j = 0; j < cnt; ++j) {
      float a[D], b[D];
      prim_box<D, KIND>(prims + (size_t)(first + j) * Prim<D, KIND>::STRIDE, a, b);
#pragma unroll
      for (int k = 0; k < D; ++k) { mn[k] = fminf(mn[k], a[k]); mx[k] = fmaxf(mx[k], b[k]); }
    }
  } else {
    const float* nd = nodes + (size_t)ref * NF;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      mn[k] = fminf(nd[4 * k + 0], nd[4 * k + 2]);
      mx[k] = fmaxf(nd[4 * k + 1], nd[4 * k + 3]);
    }
  }
}

template <int D, int KIND>
__global__ void k_refit_level(float* nodes, const float* __restrict__ prims, const uint32_t* __restrict__ level, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int NF = (D + 1) * 4;
  float* nd = nodes + (size_t)level[i] * NF;
  const uint32_t ra = __float_as_uint(nd[4 * D + 0]), rb = __float_as_uint(nd[4 * D + 1]);
  float amn[D], amx[D], bmn[D], bmx[D];
  child_box<D, KIND>(nodes, prims, ra, amn, amx);
  child_box<D, KIND>(nodes, prims, rb, bmn, bmx);
#pragma unroll
  for (int k = 0; k < D; ++k) {
    nd[4 * k + 0] = amn[k]; nd[4 * k + 1] = amx[k];
    nd[4 * k + 2] = bmn[k]; nd[4 * k + 3] = bmx[k];
  }
}

template <int D, int KIND>
__global__ void k_refit_root(const float* nodes, const float* __restrict__ prims, float* root) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t ref = __float_as_uint(root[2 * D]);
  float mn[D], mx[D];
  child_box<D, KIND>(nodes, prims, ref, mn, mx);
#pragma unroll
  for (int k = 0; k < D; ++k) { root[k] = mn[k]; root[D + k] = mx[k]; }
}

// Interior surface sum (tree.rs:104-110) on the device: Aabb::half_surface
// (aabb.rs:84-97) of every interior node of the reference tree, each term in
// f32 in the reference's operation order, summed in f64 by a fixed-order
// reduction (the same value on every call for the same bounds; the reference's
// f32 running sum in node order can differ from it in its last bits).  The
// interior nodes are the root when its ref is a pair node, and every child slot
// of a pair node whose ref is a pair node that is not a synthetic chain node
// (a synthetic chain stands for one over-full reference leaf).
template <int D>
__device__ __forceinline__ float half_surface_of(const float* lo, const float* hi, int stride) {
  float sum = 0.0f;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    float prod = 1.0f;
#pragma unroll
    for (int c = 0; c < D; ++c)
      if (c != a) prod = prod * (hi[c * stride] - lo[c * stride]);
    sum = sum + prod;
  }
  return sum;
}

template <int D>
__device__ __forceinline__ bool interior_ref(const float* nodes, uint32_t r) {
  constexpr int NF = (D + 1) * 4;
  return !(r & LEAF_BIT) && __float_as_uint(nodes[(size_t)r * NF + 4 * D + 3]) == 0u;
}

constexpr int SURF_BLOCKS = 512;

template <int D>
__global__ __launch_bounds__(256) void k_surface_partials(const float* __restrict__ nodes, uint32_t n_pairs, double* part) {
  constexpr int NF = (D + 1) * 4;
  __shared__ double red[256];
  double s = 0.0;
  for (uint32_t p = blockIdx.x * 256u + threadIdx.x; p < n_pairs; p += gridDim.x * 256u) {
    const float* nd = nodes + (size_t)p * NF;
#pragma unroll
    for (int c = 0; c < 2; ++c)
      if (interior_ref<D>(nodes, __float_as_uint(nd[4 * D + c]))) s += (double)half_surface_of<D>(nd + 2 * c, nd + 2 * c + 1, 4);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// The partials in block order plus the root's term.
template <int D>
__global__ void k_surface_total(const double* __restrict__ part, uint32_t n_part, const float* __restrict__ nodes,
                                const float* __restrict__ root, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (uint32_t i = 0; i < n_part; ++i) s += part[i];
  if (interior_ref<D>(nodes, __float_as_uint(root[2 * D]))) s += (double)half_surface_of<D>(root, root + D, 1);
  *out = s;
}

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

template <class T>
static int scan_into(sq_ctx* c, const T* in, uint64_t n, uint64_t* out, std::vector<std::unique_ptr<DevBuf>>& temps) {
  const uint64_t tiles = (n + 1 + 4095) / 4096;
  temps.emplace_back(new DevBuf());
  DevBuf& sums = *temps.back();
  int rc = sums.ensure(tiles * 8);
  if (rc) return rc;
  k_scan_tile<T><<<(uint32_t)tiles, 256, 0, c->stream>>>(in, n, out, sums.as<uint64_t>());
  HIP_TRY(hipGetLastError());
  if (tiles > 1) {
    temps.emplace_back(new DevBuf());
    DevBuf& toff = *temps.back();
    if ((rc = toff.ensure((tiles + 1) * 8))) return rc;
    if ((rc = scan_into<uint64_t>(c, sums.as<uint64_t>(), tiles, toff.as<uint64_t>(), temps))) return rc;
    k_scan_add<<<grid_for(n + 1, 256), 256, 0, c->stream>>>(out, n + 1, toff.as<uint64_t>());
    HIP_TRY(hipGetLastError());
  }
  return SQ_OK;
}

int scan_counts(sq_ctx* c, const uint32_t* in, uint64_t n, uint64_t* out, std::vector<std::unique_ptr<DevBuf>>& temps) {
  return scan_into<uint32_t>(c, in, n, out, temps);
}

void ctx_release(sq_ctx* c) {
  if (--c->refs > 0) return;
  (void)hipSetDevice(c->device);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->own) (void)hipStreamDestroy(c->own);
  delete c;
}

int fill_miss(sq_ctx* c, uint64_t n, int dim, NearestOut out) {
  k_fill_miss<<<grid_for(n, 256), 256, 0, c->stream>>>((uint32_t)n, dim, out);
  HIP_TRY(hipGetLastError());
  return SQ_OK;
}

// Per-query sort of keys[offsets[i] .. offsets[i+1]) (lists over 64 entries
// in one block each), then unpack into out_leaf / out_toi.
static int csr_sort_unpack(sq_ctx* c, uint64_t n, const uint64_t* offsets, uint64_t tot, int32_t* out_leaf, float* out_toi,
                           bool has_t) {
  int rc;
  if ((rc = c->keys2.ensure(tot * 8))) return rc;
  if ((rc = c->scratch.ensure((n + 1) * 4 + 64))) return rc;
  uint64_t* keys = c->keys.as<uint64_t>();
  uint32_t* n_large = c->scratch.as<uint32_t>();
  uint32_t* large = n_large + 16;
  HIP_TRY(hipMemsetAsync(n_large, 0, 4, c->stream));
  k_sort_small<<<grid_for(n, 256), 256, 0, c->stream>>>(keys, offsets, (uint32_t)n, large, n_large);
  HIP_TRY(hipGetLastError());
  uint32_t h_large = 0;
  HIP_TRY(hipMemcpyAsync(&h_large, n_large, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (h_large) {
    k_sort_large<<<h_large, 256, 0, c->stream>>>(keys, c->keys2.as<uint64_t>(), offsets, large);
    HIP_TRY(hipGetLastError());
  }
  k_unpack<<<grid_for(tot, 256), 256, 0, c->stream>>>(keys, tot, out_leaf, out_toi, has_t ? 1 : 0);
  HIP_TRY(hipGetLastError());
  return SQ_OK;
}

// count -> scan -> capacity check -> fill -> per-query sort -> unpack
int csr_run(sq_ctx* c, uint64_t n, uint64_t* offsets, int32_t* out_leaf, float* out_toi, bool has_t,
            uint64_t capacity, uint64_t* total, const std::function<int(bool, uint64_t*)>& launch) {
  int rc;
  if ((rc = launch(false, nullptr))) return rc;
  std::vector<std::unique_ptr<DevBuf>> temps;
  if ((rc = scan_into<uint32_t>(c, c->counts.as<uint32_t>(), n, offsets, temps))) return rc;
  HIP_TRY(hipMemcpyAsync(total, offsets + n, 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const uint64_t tot = *total;
  if (tot > capacity) return fail(SQ_ERR_CAPACITY, "result needs " + std::to_string(tot) + " entries, capacity is " + std::to_string(capacity));
  if (!tot) return SQ_OK;
  if (!out_leaf || (has_t && !out_toi)) return fail(SQ_ERR_INVALID, "output arrays are NULL");
  if ((rc = c->keys.ensure(tot * 8))) return rc;
  if ((rc = launch(true, c->keys.as<uint64_t>()))) return rc;
  return csr_sort_unpack(c, n, offsets, tot, out_leaf, out_toi, has_t);
}

// Entries the first staged walk of a call makes room for (beyond what the
// staging area already holds): 16M entries, 256 MiB of blocks.
static constexpr uint64_t STAGE_FIRST = 1ull << 24;

// Pooled collect: launch(want_entries, want_blocks, walk) walks every query once,
// appending its entries to staging blocks sized for at least want_entries and,
// when it splits queries, writing sub-walks to item blocks (at least
// want_blocks per expand stage; 0: the default); it leaves per-query counts and
// the words counter[1] staging blocks claimed, [2] staging overflow, [4] item
// overflow, [6] and [8] item blocks claimed by the two expand stages.  Then scan -> capacity check -> scatter -> per-query sort ->
// unpack.  A walk whose sub-walks did not fit (its counts are partial) runs
// again with room for all of them; one whose entries did not fit runs again
// with room for the total, once the total is known to fit the capacity.  The
// buffers keep their size for later calls.
int csr_run_staged(sq_ctx* c, uint64_t n, uint64_t* offsets, int32_t* out_leaf, float* out_toi, bool has_t,
                   uint64_t capacity, uint64_t* total, const std::function<int(uint64_t, uint64_t, StagedWalk&)>& launch) {
  StagedWalk walk{};
  int rc;
  uint64_t want_e = std::min<uint64_t>(capacity, STAGE_FIRST), want_i = 0;
  uint32_t words[16];
  std::vector<std::unique_ptr<DevBuf>> temps;
  for (int pass = 0;; ++pass) {
    if (pass == 4) return fail(SQ_ERR_HIP, "collect staging overflowed");
    if ((rc = launch(want_e, want_i, walk))) return rc;
    if ((rc = scan_into<uint32_t>(c, c->counts.as<uint32_t>(), n, offsets, temps))) return rc;
    HIP_TRY(hipMemcpyAsync(total, offsets + n, 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(words, c->counter.as<uint32_t>(), sizeof(words), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (words[4]) {
      // blocks claimed vary with the waves' interleaving: leave a quarter more
      want_i = std::max(words[6], words[8]);
      want_i += want_i / 4 + 1024;
      if (want_i >= 0xFFFFFFF0ull / ITEM_BLK) return fail(SQ_ERR_INVALID, "too many sub-walks for one call; lower split_depth");
      continue;
    }
    const uint64_t tot = *total;
    if (tot > capacity) return fail(SQ_ERR_CAPACITY, "result needs " + std::to_string(tot) + " entries, capacity is " + std::to_string(capacity));
    if (!tot) return SQ_OK;
    if (!out_leaf || (has_t && !out_toi)) return fail(SQ_ERR_INVALID, "output arrays are NULL");
    if (words[2]) {
      want_e = tot;
      continue;
    }
    break;
  }
  const uint64_t tot = *total;
  if ((rc = c->keys.ensure(tot * 8))) return rc;
  const uint64_t n_blocks = words[1];
  if (n_blocks > c->stage.bytes / (STAGE_BLK * sizeof(StageEntry)) || n_blocks > c->stage_used.bytes / 4)
    return fail(SQ_ERR_HIP, "collect staging block count out of range");
  k_stage_scatter<<<grid_for(n_blocks * STAGE_BLK, 256), 256, 0, c->stream>>>(
      c->stage.as<StageEntry>(), c->stage_used.as<uint32_t>(), n_blocks, offsets, walk.items, walk.item_base,
      c->keys.as<uint64_t>());
  HIP_TRY(hipGetLastError());
  return csr_sort_unpack(c, n, offsets, tot, out_leaf, out_toi, has_t);
}

extern "C" {

const char* sq_last_error(void) { return g_err.c_str(); }
int sq_version(void) { return 100; }

int sq_ctx_create(int device, sq_ctx** out) {
  if (!out) return fail(SQ_ERR_INVALID, "out is NULL");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(SQ_ERR_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= count) return fail(SQ_ERR_NO_DEVICE, "device index out of range");
  HIP_TRY(hipSetDevice(device));
  auto c = std::make_unique<sq_ctx>();
  c->device = device;
  HIP_TRY(hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device));
  HIP_TRY(hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking));
  c->stream = c->own;
  if (int rc = c->pool_ctr.ensure(256)) return rc;
  HIP_TRY(hipMemsetAsync(c->pool_ctr.p, 0, 256, c->own));
  HIP_TRY(hipStreamSynchronize(c->own));
  if (int rc = scratch_init(device)) return rc;   // sq_raycast_nearest_noalloc's memory on this device
  *out = c.release();
  return SQ_OK;
}

int sq_ctx_destroy(sq_ctx* c) {
  if (!c) return SQ_OK;
  (void)hipStreamSynchronize(c->stream);
  ctx_release(c);
  return SQ_OK;
}

int sq_ctx_set_stream(sq_ctx* c, void* stream) {
  if (!c) return fail(SQ_ERR_INVALID, "ctx is NULL");
  c->stream = (hipStream_t)stream;
  return SQ_OK;
}

int sq_ctx_reset_stream(sq_ctx* c) {
  if (!c) return fail(SQ_ERR_INVALID, "ctx is NULL");
  c->stream = c->own;
  return SQ_OK;
}

int sq_ctx_synchronize(sq_ctx* c) {
  if (!c) return fail(SQ_ERR_INVALID, "ctx is NULL");
  HIP_TRY(hipStreamSynchronize(c->stream));
  return SQ_OK;
}

void* sq_ctx_stream(sq_ctx* c) { return c ? (void*)c->stream : nullptr; }

int sq_ctx_set_tuning(sq_ctx* c, const sq_tuning* t) {
  if (!c || !t) return fail(SQ_ERR_INVALID, "NULL argument");
  const int block = t->block ? t->block : 128;
  if (block < 64 || block > 256 || block % 64) return fail(SQ_ERR_INVALID, "block must be 64, 128, 192 or 256");
  if (t->rays_per_lane < 0 || t->rays_per_lane > 4096) return fail(SQ_ERR_INVALID, "rays_per_lane out of range");
  const int S = t->lds_stack ? t->lds_stack : 8;
  if (S < 2 || S > 64 || (S & (S - 1))) return fail(SQ_ERR_INVALID, "lds_stack must be a power of two in [2, 64]");
  if (t->ray_order != 0 && t->ray_order != 1) return fail(SQ_ERR_INVALID, "ray_order must be 0 or 1");
  if (t->descend_min < 0 || t->descend_min > 64) return fail(SQ_ERR_INVALID, "descend_min must be in [0, 64]");
  if (t->split_depth < -1 || t->split_depth > 24) return fail(SQ_ERR_INVALID, "split_depth must be in [-1, 24]");
  if (t->wide < 0 || t->wide > 2) return fail(SQ_ERR_INVALID, "wide must be 0, 1 or 2");
  if (t->packet != 0 && t->packet != 1) return fail(SQ_ERR_INVALID, "packet must be 0 or 1");
  if (t->lookahead != 0 && t->lookahead != 1) return fail(SQ_ERR_INVALID, "lookahead must be 0 or 1");
  if (t->lookahead_pause < 0 || t->lookahead_pause > 64) return fail(SQ_ERR_INVALID, "lookahead_pause must be in [0, 64]");
  c->block = block;
  c->rays_per_lane = t->rays_per_lane;
  c->lds_stack = S;
  c->ray_order = t->ray_order;
  c->descend_min = t->descend_min;
  c->split_depth = t->split_depth;
  c->wide = t->wide;
  c->packet = t->packet;
  c->lookahead = t->lookahead;
  c->lookahead_pause = t->lookahead_pause;
  return SQ_OK;
}

int sq_ctx_set_hit_mode(sq_ctx* c, int mode) {
  if (!c) return fail(SQ_ERR_INVALID, "ctx is NULL");
  if (mode != SQ_HITS_REFERENCE && mode != SQ_HITS_GUARDED) return fail(SQ_ERR_INVALID, "hit mode must be SQ_HITS_REFERENCE or SQ_HITS_GUARDED");
  c->hit_mode = mode;
  return SQ_OK;
}

int sq_ctx_get_hit_mode(const sq_ctx* c) { return c ? c->hit_mode : -1; }

int sq_ctx_get_tuning(const sq_ctx* c, sq_tuning* t) {
  if (!c || !t) return fail(SQ_ERR_INVALID, "NULL argument");
  t->block = c->block;
  t->rays_per_lane = c->rays_per_lane;
  t->lds_stack = c->lds_stack;
  t->ray_order = c->ray_order;
  t->descend_min = c->descend_min;
  t->split_depth = c->split_depth;
  t->wide = c->wide;
  t->packet = c->packet;
  t->lookahead = c->lookahead;
  t->lookahead_pause = c->lookahead_pause;
  return SQ_OK;
}

int sq_ctx_set_block(sq_ctx* c, int block) {
  if (!c) return fail(SQ_ERR_INVALID, "ctx is NULL");
  if (block == 0) block = 128;
  if (block < 64 || block > 1024 || block % 64) return fail(SQ_ERR_INVALID, "block must be a multiple of 64 in [64, 1024]");
  if (c->rays_per_lane && block > 256) return fail(SQ_ERR_INVALID, "the pooled kernels take blocks of at most 256");
  c->block = block;
  return SQ_OK;
}

int sq_ctx_last_kernel_ms(sq_ctx* c, float* ms) {
  if (!c || !ms) return fail(SQ_ERR_INVALID, "NULL argument");
  if (!c->timed) return fail(SQ_ERR_INVALID, "no timed kernel yet");
  HIP_TRY(hipEventSynchronize(c->ev1));
  HIP_TRY(hipEventElapsedTime(ms, c->ev0, c->ev1));
  return SQ_OK;
}

struct sq_tree {
  sqh::RefTree t;
};
}  // extern "C"

void tree_stats_of(const sqh::RefTree& t, sq_tree_stats& st) {
  st.node_count = t.node_count();
  sqh::tree_structural(t, &st.leaf_count, &st.prim_count, &st.max_depth);
  // size_of::<Node<D>>() = 8D + 12 bytes, plus u32 prim indices
  st.bytes = st.node_count * (uint64_t)(8 * t.dim + 12) + t.prim_indices.size() * 4;
  st.quality = sqh::tree_interior_surface_sum(t);
}

extern "C" {


static int check_dim_kind(int dim, int kind) {
  return dispatch_dk(dim, kind, [&]<int D, int K>() { return SQ_OK; });
}

int sq_tree_build(int dim, int kind, const float* prims, uint64_t n, sq_tree** out) {
  if (!out || (n && !prims)) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = nullptr;
  int rc = check_dim_kind(dim, kind);
  if (rc) return rc;
  if (n > sqh::LEAF_FIRST_MASK) return fail(SQ_ERR_INVALID, "at most 2^27 primitives per tree");
  auto t = std::make_unique<sq_tree>();
  std::vector<float> boxes;
  sqh::prim_boxes(dim, kind, prims, n, boxes);
  sqh::build_tree(dim, boxes.data(), (uint32_t)n, t->t);
  *out = t.release();
  return SQ_OK;
}

int sq_tree_from_nodes(int dim, const float* node_bounds, const uint32_t* child_a, const uint32_t* child_b,
                       const uint32_t* prim_count, uint64_t n_nodes, const uint32_t* prim_indices, uint64_t n,
                       sq_tree** out) {
  if (!out) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (dim < 1 || dim > sqh::MAX_DIM) return fail(SQ_ERR_UNSUPPORTED, "dimension out of range");
  if (n > sqh::LEAF_FIRST_MASK) return fail(SQ_ERR_INVALID, "at most 2^27 primitives per tree");
  if ((n_nodes == 0) != (n == 0)) return fail(SQ_ERR_INVALID, "an empty tree has no nodes and no primitives");
  if (n_nodes && (!node_bounds || !child_a || !child_b || !prim_count || !prim_indices)) return fail(SQ_ERR_INVALID, "NULL argument");
  auto tp = std::make_unique<sq_tree>();
  sqh::RefTree& t = tp->t;
  t.dim = dim;
  t.bounds.assign(node_bounds, node_bounds + n_nodes * 2 * dim);
  t.child_a.assign(child_a, child_a + n_nodes);
  t.child_b.assign(child_b, child_b + n_nodes);
  t.prim_count.assign(prim_count, prim_count + n_nodes);
  t.prim_indices.assign(prim_indices, prim_indices + n);
  // The device walks this topology: every link must point forward (the
  // reference's pre-order guarantee, tree.rs:35-38) and every leaf range must
  // lie inside prim_indices.
  for (uint64_t i = 0; i < n_nodes; ++i) {
    if (t.prim_count[i] == 0) {
      if (t.child_a[i] <= i || t.child_b[i] <= i || t.child_a[i] >= n_nodes || t.child_b[i] >= n_nodes)
        return fail(SQ_ERR_INVALID, "child links must point forward inside the node array");
    } else if ((uint64_t)t.child_a[i] + t.prim_count[i] > n) {
      return fail(SQ_ERR_INVALID, "leaf range outside prim_indices");
    }
  }
  for (uint64_t i = 0; i < n; ++i) if (t.prim_indices[i] >= n) return fail(SQ_ERR_INVALID, "prim index out of range");
  *out = tp.release();
  return SQ_OK;
}

int sq_tree_refit(sq_tree* tp, int kind, const float* prims) {
  if (!tp || (tp->t.prim_indices.size() && !prims)) return fail(SQ_ERR_INVALID, "NULL argument");
  sqh::RefTree& t = tp->t;
  int rc = check_dim_kind(t.dim, kind);
  if (rc) return rc;
  const int D = t.dim;
  std::vector<float> boxes;
  sqh::prim_boxes(D, kind, prims, t.prim_indices.size(), boxes);
  // tree.rs:54-75: one reverse pass, leaves fold their prims from EMPTY.
  for (size_t ni = t.node_count(); ni-- > 0;) {
    float* b = t.bounds.data() + ni * 2 * D;
    if (t.prim_count[ni] > 0) {
      for (int k = 0; k < D; ++k) { b[k] = INFINITY; b[D + k] = -INFINITY; }
      for (uint32_t j = t.child_a[ni]; j < t.child_a[ni] + t.prim_count[ni]; ++j) {
        const float* pb = boxes.data() + (size_t)t.prim_indices[j] * 2 * D;
        for (int k = 0; k < D; ++k) { b[k] = fminf(b[k], pb[k]); b[D + k] = fmaxf(b[D + k], pb[D + k]); }
      }
    } else {
      const float* ba = t.bounds.data() + (size_t)t.child_a[ni] * 2 * D;
      const float* bb = t.bounds.data() + (size_t)t.child_b[ni] * 2 * D;
      for (int k = 0; k < D; ++k) { b[k] = fminf(ba[k], bb[k]); b[D + k] = fmaxf(ba[D + k], bb[D + k]); }
    }
  }
  return SQ_OK;
}

int sq_tree_get_stats(const sq_tree* tp, sq_tree_stats* out) {
  if (!tp || !out) return fail(SQ_ERR_INVALID, "NULL argument");
  tree_stats_of(tp->t, *out);
  return SQ_OK;
}

int sq_tree_export(const sq_tree* tp, float* bounds, uint32_t* ca, uint32_t* cb, uint32_t* pc, uint32_t* idx) {
  if (!tp) return fail(SQ_ERR_INVALID, "tree is NULL");
  const sqh::RefTree& t = tp->t;
  if (bounds) std::memcpy(bounds, t.bounds.data(), t.bounds.size() * 4);
  if (ca) std::memcpy(ca, t.child_a.data(), t.child_a.size() * 4);
  if (cb) std::memcpy(cb, t.child_b.data(), t.child_b.size() * 4);
  if (pc) std::memcpy(pc, t.prim_count.data(), t.prim_count.size() * 4);
  if (idx) std::memcpy(idx, t.prim_indices.data(), t.prim_indices.size() * 4);
  return SQ_OK;
}

int sq_tree_destroy(sq_tree* tp) {
  delete tp;
  return SQ_OK;
}

int sq_bvh_upload(sq_ctx* ctx, const sq_tree* tp, int kind, const float* prims, sq_bvh** out) {
  if (!ctx || !tp || !out) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = nullptr;
  const sqh::RefTree& t = tp->t;
  const int dim = t.dim;
  int rc = check_dim_kind(dim, kind);
  if (rc) return rc;
  const uint64_t n = t.prim_indices.size();
  if (n && !prims) return fail(SQ_ERR_INVALID, "prims is NULL");
  auto b = std::make_unique<sq_bvh>();
  b->dim = dim; b->kind = kind; b->n_prims = n;
  b->tree = t;
  try {
    sqh::pack_tree(b->tree, b->packed);
  } catch (const std::exception& e) {
    return fail(SQ_ERR_INVALID, e.what());
  }
  std::vector<float> dp;
  sqh::pack_prims(dim, kind, prims, b->tree, dp);
  hipStream_t s = ctx->stream;
  HIP_TRY(hipSetDevice(ctx->device));
  if ((rc = upload(b->nodes, b->packed.nodes.data(), b->packed.nodes.size() * 4, s))) return rc;
  if ((rc = upload(b->prims, dp.data(), dp.size() * 4, s))) return rc;
  if ((rc = upload(b->ids, b->tree.prim_indices.data(), b->tree.prim_indices.size() * 4, s))) return rc;
  if ((rc = upload(b->root, b->packed.root.data(), b->packed.root.size() * 4, s))) return rc;
  if ((rc = upload(b->levels, b->packed.levels.data(), b->packed.levels.size() * 4, s))) return rc;
  HIP_TRY(hipStreamSynchronize(s));
  tree_stats_of(b->tree, b->stats);
  b->ctx = ctx;
  ctx_retain(ctx);
  *out = b.release();
  return SQ_OK;
}

int sq_bvh_build(sq_ctx* ctx, int dim, int kind, const float* prims, uint64_t n, sq_bvh** out) {
  if (!ctx || !out) return fail(SQ_ERR_INVALID, "NULL argument");
  sq_tree* t = nullptr;
  int rc = sq_tree_build(dim, kind, prims, n, &t);
  if (rc) return rc;
  rc = sq_bvh_upload(ctx, t, kind, prims, out);
  sq_tree_destroy(t);
  return rc;
}

int sq_bvh_export(sq_bvh* b, float* bounds, uint32_t* ca, uint32_t* cb, uint32_t* pc, uint32_t* idx) {
  if (!b) return fail(SQ_ERR_INVALID, "bvh is NULL");
  int rc = ensure_host_tree(b);
  if (rc) return rc;
  const sqh::RefTree& t = b->tree;
  if (bounds && (rc = sq_bvh_download_bounds(b, bounds))) return rc;
  if (ca) std::memcpy(ca, t.child_a.data(), t.child_a.size() * 4);
  if (cb) std::memcpy(cb, t.child_b.data(), t.child_b.size() * 4);
  if (pc) std::memcpy(pc, t.prim_count.data(), t.prim_count.size() * 4);
  if (idx) std::memcpy(idx, t.prim_indices.data(), t.prim_indices.size() * 4);
  return SQ_OK;
}

int sq_bvh_destroy(sq_bvh* b) {
  if (!b) return SQ_OK;
  sq_ctx* c = b->ctx;
  if (c) (void)hipStreamSynchronize(c->stream);
  delete b;
  if (c) ctx_release(c);
  return SQ_OK;
}

int sq_bvh_set_layers(sq_bvh* b, const uint64_t* layers) {
  if (!b) return fail(SQ_ERR_INVALID, "bvh is NULL");
  if (!layers) {
    b->layers.~DevBuf();
    new (&b->layers) DevBuf();
    return SQ_OK;
  }
  int rc = upload(b->layers, layers, b->n_prims * 8, b->ctx->stream);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  return SQ_OK;
}

int sq_bvh_get_stats(const sq_bvh* b, sq_tree_stats* out) {
  if (!b || !out) return fail(SQ_ERR_INVALID, "NULL argument");
  int rc = ensure_host_tree(const_cast<sq_bvh*>(b));     // quality is a host sum in node order
  if (rc) return rc;
  *out = b->stats;
  return SQ_OK;
}

int sq_bvh_download_bounds(sq_bvh* b, float* bounds) {
  if (!b || !bounds) return fail(SQ_ERR_INVALID, "NULL argument");
  int rc0 = ensure_host_tree(b);
  if (rc0) return rc0;
  const int D = b->dim, NF = (D + 1) * 4;
  const size_t nn = b->tree.node_count();
  if (!nn) return SQ_OK;
  std::vector<float> nodes(b->packed.nodes.size()), root(16);
  HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  if (!nodes.empty()) HIP_TRY(hipMemcpy(nodes.data(), b->nodes.p, nodes.size() * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(root.data(), b->root.p, 16 * 4, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < nn; ++i) {
    float* o = bounds + i * 2 * D;
    const uint32_t slot = b->packed.slot_of_ref[i];
    if (slot == 0xFFFFFFFFu) {
      std::memcpy(o, root.data(), 2 * D * 4);
    } else {
      const float* nd = nodes.data() + (size_t)(slot >> 1) * NF;
      const int s = slot & 1;
      for (int k = 0; k < D; ++k) { o[k] = nd[4 * k + 2 * s]; o[D + k] = nd[4 * k + 2 * s + 1]; }
    }
  }
  return SQ_OK;
}

int sq_bvh_refit(sq_bvh* b, const float* prims_device) {
  if (!b || (b->n_prims && !prims_device)) return fail(SQ_ERR_INVALID, "NULL argument");
  if (!b->n_prims) return SQ_OK;
  sq_ctx* c = b->ctx;
  return dispatch_dk(b->dim, b->kind, [&]<int D, int K>() -> int {
    const uint32_t n = (uint32_t)b->n_prims;
    k_scatter_prims<D, K><<<grid_for(n, 256), 256, 0, c->stream>>>(prims_device, b->ids.as<uint32_t>(), b->prims.as<float>(), n);
    HIP_TRY(hipGetLastError());
    const auto& lo = b->packed.level_off;
    for (size_t l = 0; l + 1 < lo.size(); ++l) {
      const uint32_t cnt = lo[l + 1] - lo[l];
      if (!cnt) continue;
      k_refit_level<D, K><<<grid_for(cnt, 256), 256, 0, c->stream>>>(b->nodes.as<float>(), b->prims.as<float>(),
                                                                    b->levels.as<uint32_t>() + lo[l], cnt);
      HIP_TRY(hipGetLastError());
    }
    k_refit_root<D, K><<<1, 64, 0, c->stream>>>(b->nodes.as<float>(), b->prims.as<float>(), b->root.as<float>());
    HIP_TRY(hipGetLastError());
    return SQ_OK;
  });
}

// The device interior surface sum of b's current bounds into *out (host);
// synchronises the context's stream.
static int surface_sum(sq_bvh* b, double* out) {
  sq_ctx* c = b->ctx;
  *out = 0.0;
  const uint32_t n_pairs = b->packed.n_pairs;
  int rc = c->scratch.ensure((SURF_BLOCKS + 1) * sizeof(double));
  if (rc) return rc;
  double* part = c->scratch.as<double>();
  auto go = [&]<int D>() -> int {
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>(SURF_BLOCKS, (n_pairs + 255) / 256));
    k_surface_partials<D><<<blocks, 256, 0, c->stream>>>(b->nodes.as<float>(), n_pairs, part);
    HIP_TRY(hipGetLastError());
    k_surface_total<D><<<1, 64, 0, c->stream>>>(part, blocks, b->nodes.as<float>(), b->root.as<float>(), part + SURF_BLOCKS);
    HIP_TRY(hipGetLastError());
    return SQ_OK;
  };
  switch (b->dim) {
    case 2: rc = go.template operator()<2>(); break;
    case 3: rc = go.template operator()<3>(); break;
    case 4: rc = go.template operator()<4>(); break;
    case 5: rc = go.template operator()<5>(); break;
    case 6: rc = go.template operator()<6>(); break;
    default: return fail(SQ_ERR_UNSUPPORTED, "dimension not instantiated");
  }
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(out, part + SURF_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return SQ_OK;
}

int sq_bvh_surface_sum(sq_bvh* b, double* out) {
  if (!b || !out) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = 0.0;
  if (!b->n_prims) return SQ_OK;
  return surface_sum(b, out);
}

int sq_bvh_refit_auto(sq_bvh* b, const float* prims_device, float factor, int* rebuilt) {
  if (!b || !rebuilt || (b->n_prims && !prims_device)) return fail(SQ_ERR_INVALID, "NULL argument");
  *rebuilt = 0;
  if (!(factor >= 1.0f)) return fail(SQ_ERR_INVALID, "factor must be >= 1");
  if (!b->n_prims) return SQ_OK;
  int rc;
  if (!(b->auto_baseline > 0.0) && (rc = surface_sum(b, &b->auto_baseline))) return rc;   // the tree as built
  if ((rc = sq_bvh_refit(b, prims_device))) return rc;
  double q = 0.0;
  if ((rc = surface_sum(b, &q))) return rc;
  if (!(q > (double)factor * b->auto_baseline)) return SQ_OK;
  if ((rc = sq_bvh_rebuild(b, prims_device))) return rc;
  *rebuilt = 1;
  return surface_sum(b, &b->auto_baseline);
}

static int nearest_impl(sq_bvh* b, const float* q, uint64_t n, float max_toi, const float* tmax, const sq_filter* f,
                        NearestOut out, bool any, bool cast) {
  if (!b || (n && !q)) return fail(SQ_ERR_INVALID, "NULL argument");
  if (n > 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "at most 2^32-1 queries per call");
  if (!n) return SQ_OK;
  sq_ctx* c = b->ctx;
  const int B = c->block;
  if (!b->n_prims || (f && f->layer_mask == 0) || (cast && b->kind == KIND_TRI)) {
    k_fill_miss<<<grid_for(n, 256), 256, 0, c->stream>>>((uint32_t)n, b->dim, out);
    HIP_TRY(hipGetLastError());
    return SQ_OK;
  }
  bool filt;
  Filter fd = filter_of(b, f, &filt);
  const uint32_t stack_n = b->packed.max_stack;
  SceneView s = view_of(b);
  const uint32_t nn = (uint32_t)n;
  if (c->packet && !any && !cast && !filt) {
    // rays in the batch-adaptive Morton order, one wave-coherent walk per 64
    // of them (packet.hip)
    int rc = begin_timing(c);
    if (rc) return rc;
    const uint32_t* order = nullptr;
    if (n >= 4096 && (rc = ray_order(c, b->dim, q, nn, 2 * b->dim, &order))) return rc;
    if ((rc = packet_nearest(b, q, n, max_toi, tmax, out, c->hit_mode != SQ_HITS_GUARDED, order))) return rc;
    return end_timing(c);
  }
  if (c->rays_per_lane == 0) {
    // one query per thread, whole stack in LDS
    const size_t lds = (size_t)stack_n * B * 8;
    if (lds > 160 * 1024) return fail(SQ_ERR_INVALID, "traversal stack does not fit in LDS; lower the block size");
    int rc = begin_timing(c);
    if (rc) return rc;
    rc = dispatch_dk(b->dim, b->kind, [&]<int D, int K>() -> int {
      dim3 grid(grid_for(n, B)), block(B);
      auto go = [&]<bool EXACT_>() {
#define SQ_LAUNCH(ANY_, FILT_, CAST_) \
        k_nearest<D, K, ANY_, FILT_, CAST_, EXACT_><<<grid, block, lds, c->stream>>>(s, q, nn, max_toi, tmax, fd, out, stack_n)
        if (cast) { if (filt) SQ_LAUNCH(false, true, true); else SQ_LAUNCH(false, false, true); }
        else if (any) { if (filt) SQ_LAUNCH(true, true, false); else SQ_LAUNCH(true, false, false); }
        else { if (filt) SQ_LAUNCH(false, true, false); else SQ_LAUNCH(false, false, false); }
#undef SQ_LAUNCH
      };
      // AABB hits lie inside their boxes: the one walk is exact in both modes
      if constexpr (K == KIND_AABB) go.template operator()<true>();
      else if (c->hit_mode == SQ_HITS_GUARDED) go.template operator()<false>();
      else go.template operator()<true>();
      HIP_TRY(hipGetLastError());
      return SQ_OK;
    });
    if (rc) return rc;
    return end_timing(c);
  }
  // pooled kernel, persistent: chunks of 64 * per queries (small calls halve
  // per until every resident wave has a chunk), a grid of the resident waves
  const uint32_t S = (uint32_t)c->lds_stack;
  const uint32_t per = pooled_per_lane(c, n);
  const uint64_t chunk = 64ull * per;
  const uint64_t chunks = (n + chunk - 1) / chunk;
  // 4-wide walk (nearest and any-hit without a filter; near-first over hits
  // inside their boxes: AABB leaves, or the guarded kinds; else in the
  // reference's order over a tree without synthetic pairs): up to 3 pushes
  // per record over half the levels.  wide 2 takes it for the walks in the
  // reference's order only (nearest rays over SPHERE / TRIANGLE leaves in the
  // default hit mode), where it pays most
  const bool ref_walk = !any && c->hit_mode != SQ_HITS_GUARDED && b->kind != KIND_AABB;
  const bool quad = (c->wide == 1 || (c->wide == 2 && ref_walk)) && !cast && !filt &&
                    (b->kind == KIND_AABB || c->hit_mode == SQ_HITS_GUARDED || b->packed.n_synth == 0);
  const uint32_t walk_stack = quad ? 3 * (stack_n / 2 + 1) + 1 : stack_n;
  const uint32_t spill_rows = walk_stack > S ? walk_stack - S : 0;
  const size_t lds = (size_t)S * B * 8;
  int rc;
  if ((rc = c->counter.ensure(256))) return rc;
  rc = begin_timing(c);
  if (rc) return rc;
  if (quad) {
    const uint32_t nf = (uint32_t)(b->dim + 1) * 4;
    const uint32_t n_nodes = (uint32_t)(b->nodes.bytes / ((size_t)nf * 4));
    if ((rc = b->quads.ensure((size_t)std::max<uint32_t>(n_nodes, 1) * 2 * nf * 4))) return rc;
    // (a tree whose root is a leaf has no pair nodes and no records to build)
    if (n_nodes) rc = dispatch_dk(b->dim, b->kind, [&]<int D, int K>() -> int {
      k_quads<D><<<grid_for(n_nodes, 256), 256, 0, c->stream>>>(b->nodes.as<float>(), n_nodes, b->quads.as<float>());
      HIP_TRY(hipGetLastError());
      return SQ_OK;
    });
    if (rc) return rc;
    s.quads = b->quads.as<float>();
  }
  const uint32_t* order = nullptr;
  if (c->ray_order && !cast && n >= 4096)
    if ((rc = ray_order(c, b->dim, q, nn, 2 * b->dim, &order))) return rc;
  uint32_t* counter = c->pool_ctr.as<uint32_t>();
  rc = dispatch_dk(b->dim, b->kind, [&]<int D, int K>() -> int {
    auto launch = [&]<bool ANY_, bool FILT_, bool CAST_, bool QUAD_, bool EXACT_, bool SPEC_ = false>() -> int {
      const void* kern = reinterpret_cast<const void*>(&k_nearest_pool<D, K, ANY_, FILT_, CAST_, QUAD_, EXACT_, SPEC_>);
      int per_cu = 0;
      if (int r0 = blocks_per_cu(kern, c->device, B, lds, &per_cu)) return r0;
      const uint64_t wpb = (uint64_t)B / 64;
      const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)c->cus * std::max(1, per_cu), (chunks + wpb - 1) / wpb));
      const uint64_t lanes = grid * (uint64_t)B;
      int r2 = c->spill.ensure((size_t)spill_rows * lanes * sizeof(SpillEntry) + 64);
      if (r2) return r2;
      k_nearest_pool<D, K, ANY_, FILT_, CAST_, QUAD_, EXACT_, SPEC_><<<(uint32_t)grid, B, lds, c->stream>>>(
          s, q, nn, max_toi, tmax, fd, out, (uint32_t)chunk, counter, S, c->spill.as<SpillEntry>(), (uint32_t)lanes, order,
          (uint32_t)(SPEC_ ? c->lookahead_pause : c->descend_min));
      HIP_TRY(hipGetLastError());
      return SQ_OK;
    };
    auto pick = [&]<bool EXACT_>() -> int {
      // the look-ahead form of the walks in the reference's order
      if constexpr (EXACT_ && K != KIND_AABB) {
        if (c->lookahead && !any && !cast) {
          if (quad) return launch.template operator()<false, false, false, true, true, true>();
          return filt ? launch.template operator()<false, true, false, false, true, true>()
                      : launch.template operator()<false, false, false, false, true, true>();
        }
      }
      if (quad) return any ? launch.template operator()<true, false, false, true, EXACT_>()
                           : launch.template operator()<false, false, false, true, EXACT_>();
      if (cast) return filt ? launch.template operator()<false, true, true, false, EXACT_>()
                            : launch.template operator()<false, false, true, false, EXACT_>();
      if (any) return filt ? launch.template operator()<true, true, false, false, EXACT_>()
                           : launch.template operator()<true, false, false, false, EXACT_>();
      return filt ? launch.template operator()<false, true, false, false, EXACT_>()
                  : launch.template operator()<false, false, false, false, EXACT_>();
    };
    // AABB hits lie inside their boxes: the one walk is exact in both modes
    if constexpr (K == KIND_AABB) return pick.template operator()<true>();
    else return c->hit_mode == SQ_HITS_GUARDED ? pick.template operator()<false>() : pick.template operator()<true>();
  });
  if (rc) return rc;
  return end_timing(c);
}

int sq_raycast_nearest(sq_bvh* b, const float* rays, uint64_t n, float max_toi, const float* tmax, const sq_filter* f,
                       int32_t* out_leaf, float* out_toi, float* out_point, float* out_normal, int32_t* out_sub) {
  if (n && (!out_leaf || !out_toi)) return fail(SQ_ERR_INVALID, "out_leaf and out_toi are required");
  NearestOut o{out_leaf, out_toi, out_point, out_normal, out_sub, nullptr};
  return nearest_impl(b, rays, n, max_toi, tmax, f, o, false, false);
}

// sq_raycast_nearest without allocating: the pooled walk's counters and spill
// columns live in the device's scratch array.  The ring stack kept in LDS is
// the context's lds_stack, doubled (up to 64 entries) until the resident grid's
// spill fits the array; past that the grid shrinks to what fits.  No timing
// events (last_kernel_ms keeps the previous call's time).
int sq_raycast_nearest_noalloc(sq_bvh* b, const float* rays, uint64_t n, float max_toi, const float* tmax,
                               const sq_filter* f, int32_t* out_leaf, float* out_toi, float* out_point, float* out_normal,
                               int32_t* out_sub) {
  if (!b || (n && (!rays || !out_leaf || !out_toi))) return fail(SQ_ERR_INVALID, "NULL argument");
  if (n > 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "at most 2^32-1 queries per call");
  if (!n) return SQ_OK;
  sq_ctx* c = b->ctx;
  NearestOut out{out_leaf, out_toi, out_point, out_normal, out_sub, nullptr};
  if (!b->n_prims || (f && f->layer_mask == 0)) {
    k_fill_miss<<<grid_for(n, 256), 256, 0, c->stream>>>((uint32_t)n, b->dim, out);
    HIP_TRY(hipGetLastError());
    return SQ_OK;
  }
  void* base = (c->device >= 0 && c->device < SQ_MAX_DEVICES) ? g_scratch[c->device] : nullptr;
  if (!base) return fail(SQ_ERR_HIP, "no scratch address for this device (context not created through sq_ctx_create)");
  bool filt;
  Filter fd = filter_of(b, f, &filt);
  const uint32_t stack_n = b->packed.max_stack;
  const SceneView s = view_of(b);
  const uint32_t nn = (uint32_t)n;
  const int B = c->block;
  const uint32_t per = pooled_per_lane(c, n);
  const uint64_t chunk = 64ull * per;
  const uint64_t chunks = (n + chunk - 1) / chunk;
  const uint64_t wpb = (uint64_t)B / 64;
  const uint64_t spill_cap = SQ_NOALLOC_SCRATCH_BYTES - 256;
  uint32_t* counter = static_cast<uint32_t*>(base);
  SpillEntry* spill = reinterpret_cast<SpillEntry*>(static_cast<unsigned char*>(base) + 256);
  return dispatch_dk(b->dim, b->kind, [&]<int D, int K>() -> int {
    auto launch = [&]<bool FILT_, bool EXACT_>() -> int {
      const void* kern = reinterpret_cast<const void*>(&k_nearest_pool<D, K, false, FILT_, false, false, EXACT_>);
      uint32_t S = (uint32_t)c->lds_stack;
      uint64_t grid = 0;
      for (;;) {
        int per_cu = 0;
        HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, B, (size_t)S * B * 8));
        grid = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)c->cus * std::max(1, per_cu), (chunks + wpb - 1) / wpb));
        const uint64_t rows = stack_n > S ? stack_n - S : 0;
        if (rows * grid * B * sizeof(SpillEntry) <= spill_cap) break;
        if (S < 64 && (size_t)(2 * S) * B * 8 <= 64 * 1024) { S *= 2; continue; }
        grid = spill_cap / (rows * B * sizeof(SpillEntry));
        if (!grid) return fail(SQ_ERR_UNSUPPORTED, "tree too deep for the no-allocation scratch");
        break;
      }
      const uint64_t lanes = grid * (uint64_t)B;
      k_zero_words<<<1, 64, 0, c->stream>>>(counter, 64);   // the pool counters start at zero
      HIP_TRY(hipGetLastError());
      k_nearest_pool<D, K, false, FILT_, false, false, EXACT_><<<(uint32_t)grid, B, (size_t)S * B * 8, c->stream>>>(
          s, rays, nn, max_toi, tmax, fd, out, (uint32_t)chunk, counter, S, spill, (uint32_t)lanes, nullptr,
          (uint32_t)c->descend_min);
      HIP_TRY(hipGetLastError());
      return SQ_OK;
    };
    auto pick = [&]<bool EXACT_>() -> int {
      return filt ? launch.template operator()<true, EXACT_>() : launch.template operator()<false, EXACT_>();
    };
    // AABB hits lie inside their boxes: the one walk is exact in both modes
    if constexpr (K == KIND_AABB) return pick.template operator()<true>();
    else return c->hit_mode == SQ_HITS_GUARDED ? pick.template operator()<false>() : pick.template operator()<true>();
  });
}

// Bvh::raycast_nearest_stats / raycast_all_stats (bvh.rs:61-70, 206-215):
// the reference's counters per ray, from a walk in the reference's order over
// the reference's hits (whatever the context's hit mode).
static int counted_impl(sq_bvh* b, const float* q, uint64_t n, float max_toi, const float* tmax, const sq_filter* f, bool all,
                        int32_t* out_leaf, float* out_toi, uint32_t* out_stats) {
  if (!b || !out_stats || (n && !q)) return fail(SQ_ERR_INVALID, "NULL argument");
  if (!all && n && (!out_leaf || !out_toi)) return fail(SQ_ERR_INVALID, "out_leaf and out_toi are required");
  if (n > 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "at most 2^32-1 queries per call");
  if (!n) return SQ_OK;
  sq_ctx* c = b->ctx;
  if (b->kind != KIND_AABB && b->kind != KIND_SPHERE && b->kind != KIND_TRI)
    return fail(SQ_ERR_UNSUPPORTED, "counters are kept for the built-in primitive kinds");
  if (!b->n_prims) {
    // an empty Bvh answers before its walk: no counts (bvh.rs:80-82)
    k_zero_u32<<<grid_for(4 * n, 256), 256, 0, c->stream>>>(out_stats, 4 * n);
    HIP_TRY(hipGetLastError());
    if (!all) {
      k_fill_miss<<<grid_for(n, 256), 256, 0, c->stream>>>((uint32_t)n, b->dim,
                                                          NearestOut{out_leaf, out_toi, nullptr, nullptr, nullptr, nullptr});
      HIP_TRY(hipGetLastError());
    }
    return SQ_OK;
  }
  bool filt;
  Filter fd = filter_of(b, f, &filt);
  const uint32_t stack_n = b->packed.max_stack;
  const uint32_t lanes = (uint32_t)std::min<uint64_t>((n + 255) / 256 * 256, 1u << 18);
  int rc = c->spill.ensure((size_t)stack_n * lanes * 8 + 64);
  if (rc) return rc;
  if ((rc = begin_timing(c))) return rc;
  SceneView s = view_of(b);
  rc = dispatch_dk(b->dim, b->kind, [&]<int D, int K>() -> int {
#define SQ_LAUNCH(ALL_, FILT_) \
    k_counted<D, K, ALL_, FILT_><<<lanes / 256, 256, 0, c->stream>>>(s, q, (uint32_t)n, max_toi, tmax, fd, out_leaf, out_toi, \
                                                                    out_stats, c->spill.as<uint32_t>(), stack_n)
    if (all) { if (filt) SQ_LAUNCH(true, true); else SQ_LAUNCH(true, false); }
    else { if (filt) SQ_LAUNCH(false, true); else SQ_LAUNCH(false, false); }
#undef SQ_LAUNCH
    HIP_TRY(hipGetLastError());
    return SQ_OK;
  });
  if (rc) return rc;
  return end_timing(c);
}

int sq_raycast_nearest_stats(sq_bvh* b, const float* rays, uint64_t n, float max_toi, const float* tmax, const sq_filter* f,
                             int32_t* out_leaf, float* out_toi, uint32_t* out_stats) {
  return counted_impl(b, rays, n, max_toi, tmax, f, false, out_leaf, out_toi, out_stats);
}

int sq_raycast_all_stats(sq_bvh* b, const float* rays, uint64_t n, float max_toi, const sq_filter* f, uint32_t* out_stats) {
  return counted_impl(b, rays, n, max_toi, nullptr, f, true, nullptr, nullptr, out_stats);
}

int sq_raycast_any(sq_bvh* b, const float* rays, uint64_t n, float max_toi, const float* tmax, const sq_filter* f, uint8_t* out_hit) {
  if (n && !out_hit) return fail(SQ_ERR_INVALID, "out_hit is required");
  NearestOut o{nullptr, nullptr, nullptr, nullptr, nullptr, out_hit};
  return nearest_impl(b, rays, n, max_toi, tmax, f, o, true, false);
}

int sq_shapecast_nearest(sq_bvh* b, const float* casts, uint64_t n, float max_toi, const float* tmax, const sq_filter* f,
                         int32_t* out_leaf, float* out_toi, float* out_point) {
  if (n && (!out_leaf || !out_toi)) return fail(SQ_ERR_INVALID, "out_leaf and out_toi are required");
  NearestOut o{out_leaf, out_toi, out_point, nullptr, nullptr, nullptr};
  return nearest_impl(b, casts, n, max_toi, tmax, f, o, false, true);
}

// count -> scan -> (check capacity) -> fill -> segment sort -> unpack
static int collect_impl(sq_bvh* b, const float* q, uint64_t n, float max_toi, const float* tmax, const sq_filter* f, int mode,
                        uint64_t* offsets, int32_t* out_leaf, float* out_toi, uint64_t capacity, uint64_t* total) {
  if (!b || !offsets || !total || (n && !q)) return fail(SQ_ERR_INVALID, "NULL argument");
  if (n > 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "at most 2^32-1 queries per call");
  sq_ctx* c = b->ctx;
  *total = 0;
  if (!n) {
    HIP_TRY(hipMemsetAsync(offsets, 0, 8, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SQ_OK;
  }
  const bool empty = !b->n_prims || (f && f->layer_mask == 0) || (mode == 3 && b->kind == KIND_TRI);
  int rc;
  if ((rc = c->counts.ensure(n * 4))) return rc;
  uint32_t* counts = c->counts.as<uint32_t>();
  if (empty) {
    HIP_TRY(hipMemsetAsync(offsets, 0, (n + 1) * 8, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SQ_OK;
  }
  bool filt;
  Filter fd = filter_of(b, f, &filt);
  const int B = c->block;
  const uint32_t stack_n = b->packed.max_stack;
  const size_t lds = (size_t)stack_n * B * 4;
  SceneView s = view_of(b);
  const uint32_t nn = (uint32_t)n;
  if (c->rays_per_lane > 0) {
    // pooled: one persistent walk appends to staging blocks (csr_run_staged)
    const uint32_t S = (uint32_t)c->lds_stack;
    const uint32_t per = pooled_per_lane(c, n);
    const uint64_t chunk = 64ull * per;
    const uint64_t chunks = (n + chunk - 1) / chunk;
    const uint32_t spill_rows = stack_n > S ? stack_n - S : 0;
    if ((rc = c->counter.ensure(256))) return rc;
    const uint32_t levels = split_levels(c, n, b->stats.max_depth, mode == 2);
    // deep splits expand in two stages: queries one per thread down l1 levels,
    // then those sub-walks, many and short, one per thread down the rest
    const uint32_t l1 = levels >= 8 ? levels / 2 : levels, l2 = levels - l1;
    auto staged = [&](uint64_t want, uint64_t want_blocks, StagedWalk& walk) -> int {
      // item blocks: room for every possible sub-walk (up to 8M of them, more
      // when a walk needed it) at the worst fill of a retired block (over
      // ITEM_BLK - 64), plus one open block per expanding wave
      auto item_blocks = [&](uint32_t lv, uint64_t waves) -> uint64_t {
        const uint64_t possible = std::min<uint64_t>(n << std::min<uint32_t>(lv, 30), 1ull << 23);
        return std::max<uint64_t>(possible / (ITEM_BLK - 64) + 1 + waves, want_blocks);
      };
      if (levels) {
        const uint64_t nb1 = item_blocks(l1, (uint64_t)grid_for(n, 256) * 4);
        int r0 = c->items.ensure(nb1 * ITEM_BLK * sizeof(WorkItem));
        if (!r0) r0 = c->items_used.ensure(nb1 * 4);
        if (!r0 && l2) {
          const uint64_t nb2 = item_blocks(levels, (uint64_t)c->cus * 8 * 4);
          r0 = c->items2.ensure(nb2 * ITEM_BLK * sizeof(WorkItem));
          if (!r0) r0 = c->items2_used.ensure(nb2 * 4);
        }
        // per sub-walk entry counts, then bases, for the stage the walk takes
        if (!r0) r0 = c->items_cnt.ensure((l2 ? c->items2.bytes : c->items.bytes) / sizeof(WorkItem) * 4);
        if (r0) return r0;
      }
      auto max_blocks_of = [](const DevBuf& it, const DevBuf& used) -> uint32_t {
        // sub-walk positions are 32-bit
        return (uint32_t)std::min<uint64_t>({it.bytes / (ITEM_BLK * sizeof(WorkItem)), used.bytes / 4, 0xFFFFFFFFull / ITEM_BLK});
      };
      const uint32_t mb1 = max_blocks_of(c->items, c->items_used), mb2 = max_blocks_of(c->items2, c->items2_used);
      return dispatch_dk(b->dim, b->kind, [&]<int D, int K>() -> int {
        auto go = [&]<int MODE_, bool FILT_, bool ITEMS_>() -> int {
          const void* kern = reinterpret_cast<const void*>(&k_collect_pool<D, K, MODE_, FILT_, ITEMS_>);
          const size_t lds_p = (size_t)S * B * 8 + (size_t)(B / 64) * 8;
          int per_cu = 0;
          if (int r0 = blocks_per_cu(kern, c->device, B, lds_p, &per_cu)) return r0;
          const uint64_t wpb = (uint64_t)B / 64;
          const uint64_t resident = (uint64_t)c->cus * std::max(1, per_cu);
          // item walks: the item count is known on the device only, so the grid is every resident wave
          const uint64_t grid = ITEMS_ ? resident : std::max<uint64_t>(1, std::min<uint64_t>(resident, (chunks + wpb - 1) / wpb));
          const uint64_t lanes = grid * (uint64_t)B;
          int r2 = c->spill.ensure((size_t)spill_rows * lanes * sizeof(SpillEntry) + 64);
          if (r2) return r2;
          // retired blocks hold over STAGE_BLK - 64 entries; each wave has one more open
          const uint64_t blocks = std::min<uint64_t>(want / (STAGE_BLK - 64) + 1 + grid * wpb, 0xFFFFFFF0ull);
          if ((r2 = c->stage.ensure(blocks * STAGE_BLK * sizeof(StageEntry)))) return r2;
          if ((r2 = c->stage_used.ensure(blocks * 4))) return r2;
          const uint64_t max_blocks = std::min<uint64_t>({c->stage.bytes / (STAGE_BLK * sizeof(StageEntry)),
                                                          c->stage_used.bytes / 4, 0xFFFFFFF0ull});
          uint32_t* counter = c->counter.as<uint32_t>();
          k_zero_words<<<1, 64, 0, c->stream>>>(counter, 16);
          HIP_TRY(hipGetLastError());
          const WorkItem* walk_items = c->items.as<WorkItem>();
          const uint32_t* walk_used = c->items_used.as<uint32_t>();
          const uint32_t* walk_blocks = counter + 6;
          uint32_t walk_max = mb1;
          if (ITEMS_) {
            k_collect_expand<D, K, MODE_, FILT_, false><<<grid_for(n, 256), 256, (size_t)l1 * 256 * 8, c->stream>>>(
                s, q, nn, max_toi, tmax, l1, nullptr, nullptr, nullptr, 0, c->items.as<WorkItem>(),
                c->items_used.as<uint32_t>(), counter + 6, mb1, counter + 4, counts);
            HIP_TRY(hipGetLastError());
            if (l2) {
              k_collect_expand<D, K, MODE_, FILT_, true><<<(uint32_t)(c->cus * 8), 256, (size_t)l2 * 256 * 8, c->stream>>>(
                  s, q, nn, max_toi, tmax, l2, c->items.as<WorkItem>(), c->items_used.as<uint32_t>(), counter + 6, mb1,
                  c->items2.as<WorkItem>(), c->items2_used.as<uint32_t>(), counter + 8, mb2, counter + 4, counts);
              HIP_TRY(hipGetLastError());
              walk_items = c->items2.as<WorkItem>();
              walk_used = c->items2_used.as<uint32_t>();
              walk_blocks = counter + 8;
              walk_max = mb2;
            }
          }
          StageOut o{c->stage.as<StageEntry>(), c->stage_used.as<uint32_t>(), counter + 1, counter + 2,
                     (uint32_t)max_blocks, counts};
          k_collect_pool<D, K, MODE_, FILT_, ITEMS_><<<(uint32_t)grid, B, lds_p, c->stream>>>(
              s, q, nn, max_toi, tmax, fd, (uint32_t)chunk, counter, S,
              c->spill.as<SpillEntry>(), (uint32_t)lanes, (uint32_t)c->descend_min, o, walk_items, walk_used, walk_blocks,
              walk_max, c->items_cnt.as<uint32_t>());
          HIP_TRY(hipGetLastError());
          walk = StagedWalk{nullptr, nullptr};
          if (ITEMS_) {
            k_item_base<<<(uint32_t)(c->cus * 4), 256, 0, c->stream>>>(walk_items, walk_used, walk_blocks, walk_max,
                                                                      c->items_cnt.as<uint32_t>(), counts);
            HIP_TRY(hipGetLastError());
            walk = StagedWalk{walk_items, c->items_cnt.as<uint32_t>()};
          }
          return SQ_OK;
        };
#define SQ_STAGED(MODE_)                                                                            \
        if (levels) return filt ? go.template operator()<MODE_, true, true>() : go.template operator()<MODE_, false, true>(); \
        return filt ? go.template operator()<MODE_, true, false>() : go.template operator()<MODE_, false, false>();
        switch (mode) {
          case 0: SQ_STAGED(0)
          case 1: SQ_STAGED(1)
          case 2: SQ_STAGED(2)
          default: SQ_STAGED(3)
        }
#undef SQ_STAGED
      });
    };
    return csr_run_staged(c, n, offsets, out_leaf, out_toi, mode != 2, capacity, total, staged);
  }
  auto launch = [&](bool fill, uint64_t* keys) -> int {
    return dispatch_dk(b->dim, b->kind, [&]<int D, int K>() -> int {
      dim3 grid(grid_for(n, B)), block(B);
#define SQ_COLLECT(MODE_)                                                                                              \
      if (fill) { if (filt) k_collect<D, K, MODE_, true, true><<<grid, block, lds, c->stream>>>(s, q, nn, max_toi, tmax, fd, counts, offsets, keys); \
                  else k_collect<D, K, MODE_, true, false><<<grid, block, lds, c->stream>>>(s, q, nn, max_toi, tmax, fd, counts, offsets, keys); } \
      else { if (filt) k_collect<D, K, MODE_, false, true><<<grid, block, lds, c->stream>>>(s, q, nn, max_toi, tmax, fd, counts, offsets, keys); \
             else k_collect<D, K, MODE_, false, false><<<grid, block, lds, c->stream>>>(s, q, nn, max_toi, tmax, fd, counts, offsets, keys); }
      switch (mode) {
        case 0: SQ_COLLECT(0) break;
        case 1: SQ_COLLECT(1) break;
        case 2: SQ_COLLECT(2) break;
        default: SQ_COLLECT(3) break;
      }
#undef SQ_COLLECT
      HIP_TRY(hipGetLastError());
      return SQ_OK;
    });
  };
  return csr_run(c, n, offsets, out_leaf, out_toi, mode != 2, capacity, total, launch);
}

int sq_raycast_all(sq_bvh* b, const float* rays, uint64_t n, float max_toi, const sq_filter* f, int crossings,
                   uint64_t* offsets, int32_t* out_leaf, float* out_toi, uint64_t capacity, uint64_t* total) {
  return collect_impl(b, rays, n, max_toi, nullptr, f, crossings ? 1 : 0, offsets, out_leaf, out_toi, capacity, total);
}

int sq_overlap(sq_bvh* b, const float* regions, uint64_t n, const sq_filter* f, uint64_t* offsets, int32_t* out_leaf,
               uint64_t capacity, uint64_t* total) {
  return collect_impl(b, regions, n, 0.0f, nullptr, f, 2, offsets, out_leaf, nullptr, capacity, total);
}

int sq_shapecast_all(sq_bvh* b, const float* casts, uint64_t n, float max_toi, const float* tmax, const sq_filter* f,
                     uint64_t* offsets, int32_t* out_leaf, float* out_toi, uint64_t capacity, uint64_t* total) {
  return collect_impl(b, casts, n, max_toi, tmax, f, 3, offsets, out_leaf, out_toi, capacity, total);
}

int sq_required_work(sq_bvh* b, const float* rays, uint64_t n, const float* tlim, uint32_t* out_nodes, uint32_t* out_prims) {
  if (!b || (n && (!rays || !tlim || !out_nodes || !out_prims))) return fail(SQ_ERR_INVALID, "NULL argument");
  if (!n) return SQ_OK;
  sq_ctx* c = b->ctx;
  if (!b->n_prims) {
    HIP_TRY(hipMemsetAsync(out_nodes, 0, n * 4, c->stream));
    HIP_TRY(hipMemsetAsync(out_prims, 0, n * 4, c->stream));
    return SQ_OK;
  }
  const int B = c->block;
  const uint32_t stack_n = b->packed.max_stack;
  const size_t lds = (size_t)stack_n * B * 4;
  SceneView s = view_of(b);
  return dispatch_dk(b->dim, b->kind, [&]<int D, int K>() -> int {
    if constexpr (K != KIND_SPHERE) {
      k_required<D><<<grid_for(n, B), B, lds, c->stream>>>(s, rays, (uint32_t)n, tlim, out_nodes, out_prims, stack_n);
      HIP_TRY(hipGetLastError());
      return SQ_OK;
    } else {
      // same kernel: the walk does not look at primitives
      k_required<D><<<grid_for(n, B), B, lds, c->stream>>>(s, rays, (uint32_t)n, tlim, out_nodes, out_prims, stack_n);
      HIP_TRY(hipGetLastError());
      return SQ_OK;
    }
  });
}

}  // extern "C"
