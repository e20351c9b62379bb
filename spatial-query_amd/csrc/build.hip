// build.hip -- the median-split build on the device.
//
// Reproduces tree::build (crates/spatial-query/src/accel/tree.rs:39-49,
// 112-163) node for node, like the host build (host_tree.cpp), but level by
// level over the whole tree at once:
//
//   * every node of a level that holds more than LEAF_SIZE primitives gets the
//     bounds of its centroids (segmented min/max over its range of the index
//     array), its longest axis (first maximum wins) and the coincidence stop
//     (extent <= f32::EPSILON keeps it a leaf), tree.rs:140-148;
//   * one stable radix sort of (segment start, total-order centroid key on the
//     segment's axis) over the whole index array sorts every splitting range
//     at once: within a range, equal keys keep their order, which is the
//     reference's stable sort_by(total_cmp) of that range (tree.rs:152-153);
//     ranges that do not split sort by their start alone and stay in place;
//   * each splitting node gets children (start, count/2) and (mid, count -
//     count/2) on the next level (tree.rs:154-157).
//
// Then subtree sizes (bottom-up) give every node its pre-order index, which
// is the reference's node order; the pair-node layout of pack_tree
// (host_tree.cpp) is written on the device, leaf chains included, and the
// device refit fills primitives and boxes.  The host copy of the reference
// tree is downloaded only when something asks for it (ensure_host_tree).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "engine_internal.h"

namespace {

constexpr uint32_t LEAF_SIZE = 4;                 // tree.rs:12
constexpr float F32_EPSILON = 1.1920929e-7f;      // f32::EPSILON
constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint32_t SMALL_SEG = 256;               // segments up to this size: one wave each
constexpr uint32_t TILE = 4096;                   // larger ones: blocks of 256 x 16 positions

// f32::total_cmp order as an unsigned key, and back.
__device__ __forceinline__ uint32_t total_ukey(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float total_unkey(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// Centroid keys of the primitives' world boxes (QueryGeometry::world_aabb of
// the built-in kinds, then Aabb::center, aabb.rs:56-58): keys[k * n + i].
template <int D, int KIND>
__global__ void k_centroid_keys(const float* __restrict__ prims, uint32_t n, uint32_t* keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int S = KIND == KIND_AABB ? 2 * D : (KIND == KIND_SPHERE ? D + 1 : 9);
  const float* p = prims + (size_t)i * S;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    float lo, hi;
    if constexpr (KIND == KIND_AABB) {
      lo = p[k];
      hi = p[D + k];
    } else if constexpr (KIND == KIND_SPHERE) {
      lo = p[k] - p[D];
      hi = p[k] + p[D];
    } else {
      lo = fminf(fminf(fminf(INFINITY, p[k]), p[3 + k]), p[6 + k]);
      hi = fmaxf(fmaxf(fmaxf(-INFINITY, p[k]), p[3 + k]), p[6 + k]);
    }
    keys[(size_t)k * n + i] = total_ukey((lo + hi) * 0.5f);
  }
}

__global__ void k_build_init(uint32_t n, uint32_t* idx, uint32_t* segstart, uint32_t* node_of, uint32_t* g_start,
                             uint32_t* g_count, uint32_t* g_level) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) {
    idx[p] = p;
    segstart[p] = 0;
    node_of[p] = p == 0 ? 0u : NONE;
  }
  if (p == 0) {
    g_start[0] = 0;
    g_count[0] = n;
    g_level[0] = 0;
  }
}

// Segment centroid bounds as total-order keys: sb[j * 2D + k] = min, [+ D] = max.
template <int D>
__global__ void k_seg_bounds_init(const uint32_t* __restrict__ count, uint32_t m, uint32_t* sb) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    sb[(size_t)j * 2 * D + k] = 0xFFFFFFFFu;
    sb[(size_t)j * 2 * D + D + k] = 0u;
  }
}

// Segments of LEAF_SIZE < count <= SMALL_SEG: one wave each, reduced in
// registers (shuffles), written directly.
template <int D>
__global__ __launch_bounds__(256) void k_seg_bounds_small(const uint32_t* __restrict__ start, const uint32_t* __restrict__ count,
                                                          uint32_t m, const uint32_t* __restrict__ idx,
                                                          const uint32_t* __restrict__ keys, uint32_t n, uint32_t* sb) {
  const uint32_t j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63u;
  if (j >= m) return;                                   // whole wave
  const uint32_t c = count[j];
  if (c <= LEAF_SIZE || c > SMALL_SEG) return;          // whole wave
  const uint32_t s = start[j];
  uint32_t mn[D], mx[D];
#pragma unroll
  for (int k = 0; k < D; ++k) { mn[k] = 0xFFFFFFFFu; mx[k] = 0u; }
  for (uint32_t o = lane; o < c; o += 64) {
    const uint32_t pi = idx[s + o];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const uint32_t v = keys[(size_t)k * n + pi];
      mn[k] = min(mn[k], v);
      mx[k] = max(mx[k], v);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      mn[k] = min(mn[k], (uint32_t)__shfl_xor((int)mn[k], off));
      mx[k] = max(mx[k], (uint32_t)__shfl_xor((int)mx[k], off));
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      sb[(size_t)j * 2 * D + k] = mn[k];
      sb[(size_t)j * 2 * D + D + k] = mx[k];
    }
  }
}

// Segments of count > SMALL_SEG: blocks of TILE positions; a block lying in
// one segment reduces in LDS and adds one atomic min/max per axis, else each
// thread flushes its runs of 16 consecutive positions.
template <int D>
__global__ __launch_bounds__(256) void k_seg_bounds_large(const uint32_t* __restrict__ segstart, const uint32_t* __restrict__ node_of,
                                                          const uint32_t* __restrict__ count, const uint32_t* __restrict__ idx,
                                                          const uint32_t* __restrict__ keys, uint32_t n, uint32_t* sb) {
  __shared__ uint32_t red[2 * D][256];
  const uint32_t p0 = blockIdx.x * TILE + threadIdx.x * 16;
  uint32_t mn[D], mx[D];
#pragma unroll
  for (int k = 0; k < D; ++k) { mn[k] = 0xFFFFFFFFu; mx[k] = 0u; }
  uint32_t run = NONE;            // node of the current run
  bool same = true;               // every position of this thread in the block's segment
  const uint32_t block_seg = segstart[min(blockIdx.x * TILE, n - 1)];
  auto flush = [&](uint32_t j) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      atomicMin(sb + (size_t)j * 2 * D + k, mn[k]);
      atomicMax(sb + (size_t)j * 2 * D + D + k, mx[k]);
      mn[k] = 0xFFFFFFFFu;
      mx[k] = 0u;
    }
  };
  // first pass: is the whole tile one large segment?
  for (uint32_t o = 0; o < 16; ++o) {
    const uint32_t p = p0 + o;
    if (p >= n) break;
    if (segstart[p] != block_seg) same = false;
  }
  const uint32_t jb = node_of[block_seg];
  const bool uniform = __syncthreads_and(same) && jb != NONE && count[jb] > SMALL_SEG;
  for (uint32_t o = 0; o < 16; ++o) {
    const uint32_t p = p0 + o;
    if (p >= n) break;
    const uint32_t j = uniform ? jb : node_of[segstart[p]];
    if (j == NONE || count[j] <= SMALL_SEG) continue;
    if (!uniform && j != run) {
      if (run != NONE) flush(run);
      run = j;
    }
    const uint32_t pi = idx[p];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const uint32_t v = keys[(size_t)k * n + pi];
      mn[k] = min(mn[k], v);
      mx[k] = max(mx[k], v);
    }
  }
  if (!uniform) {
    if (run != NONE) flush(run);
    return;
  }
#pragma unroll
  for (int k = 0; k < D; ++k) {
    red[k][threadIdx.x] = mn[k];
    red[D + k][threadIdx.x] = mx[k];
  }
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        red[k][threadIdx.x] = min(red[k][threadIdx.x], red[k][threadIdx.x + w]);
        red[D + k][threadIdx.x] = max(red[D + k][threadIdx.x], red[D + k][threadIdx.x + w]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      atomicMin(sb + (size_t)jb * 2 * D + k, red[k][0]);
      atomicMax(sb + (size_t)jb * 2 * D + D + k, red[D + k][0]);
    }
  }
}

// Split decision (tree.rs:136-148): flag[j] = 1 when node j splits, with its axis.
template <int D>
__global__ void k_decide(const uint32_t* __restrict__ count, uint32_t m, const uint32_t* __restrict__ sb, uint32_t* flag,
                         uint32_t* axis) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  uint32_t f = 0, a = 0;
  if (count[j] > LEAF_SIZE) {
    const uint32_t* b = sb + (size_t)j * 2 * D;
    float best = total_unkey(b[D]) - total_unkey(b[0]);
#pragma unroll
    for (int k = 1; k < D; ++k) {
      const float e = total_unkey(b[D + k]) - total_unkey(b[k]);
      if (e > best) { best = e; a = k; }
    }
    f = best <= F32_EPSILON ? 0u : 1u;
  }
  flag[j] = f;
  axis[j] = a;
}

// Sort keys: (segment start) << 32 | centroid key on the segment's axis for
// splitting segments, | 0 for the others.
__global__ void k_sort_keys(const uint32_t* __restrict__ segstart, const uint32_t* __restrict__ node_of,
                            const uint32_t* __restrict__ flag, const uint32_t* __restrict__ axis,
                            const uint32_t* __restrict__ idx, const uint32_t* __restrict__ keys, uint32_t n, uint64_t* out) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t s = segstart[p];
  const uint32_t j = node_of[s];
  uint32_t lo = 0;
  if (j != NONE && flag[j]) lo = keys[(size_t)axis[j] * n + idx[p]];
  out[p] = (uint64_t)s << 32 | lo;
}

// Positions in the upper half of a splitting segment move to the segment at mid.
__global__ void k_split_positions(uint32_t* segstart, const uint32_t* __restrict__ node_of, const uint32_t* __restrict__ flag,
                                  const uint32_t* __restrict__ count, uint32_t n) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t s = segstart[p];
  const uint32_t j = node_of[s];
  if (j == NONE || !flag[j]) return;
  const uint32_t mid = s + count[j] / 2;
  if (p >= mid) segstart[p] = mid;
}

// Children on the next level; node_of follows the segments' starts.
__global__ void k_children(const uint32_t* __restrict__ start, const uint32_t* __restrict__ count, uint32_t m, uint32_t base,
                           uint32_t level, const uint32_t* __restrict__ flag, const uint32_t* __restrict__ rank, uint32_t* g_left,
                           uint32_t* next_start, uint32_t* next_count, uint32_t* next_level, uint32_t base_next, uint32_t* node_of) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const uint32_t s = start[j], c = count[j];
  if (!flag[j]) {
    g_left[base + j] = NONE;
    node_of[s] = NONE;
    return;
  }
  const uint32_t r = rank[j];
  const uint32_t h = c / 2;
  g_left[base + j] = base_next + 2 * r;
  next_start[2 * r] = s;
  next_count[2 * r] = h;
  next_start[2 * r + 1] = s + h;
  next_count[2 * r + 1] = c - h;
  next_level[2 * r] = level + 1;
  next_level[2 * r + 1] = level + 1;
  node_of[s] = 2 * r;
  node_of[s + h] = 2 * r + 1;
}

__global__ void k_subtree_size(const uint32_t* __restrict__ g_left, uint32_t base, uint32_t m, uint32_t* size) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const uint32_t l = g_left[base + j];
  size[base + j] = l == NONE ? 1u : 1u + size[l] + size[l + 1];
}

__global__ void k_preorder(const uint32_t* __restrict__ g_left, const uint32_t* __restrict__ size, uint32_t base, uint32_t m,
                           uint32_t* pre) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const uint32_t g = base + j;
  if (g == 0) pre[0] = 0;
  const uint32_t l = g_left[g];
  if (l == NONE) return;
  const uint32_t pg = pre[g];
  pre[l] = pg + 1;
  pre[l + 1] = pg + 1 + size[l];
}

// The reference layout (tree.rs:20-26) in pre-order, with each node's level
// and parent.
__global__ void k_ref_nodes(const uint32_t* __restrict__ g_start, const uint32_t* __restrict__ g_count,
                            const uint32_t* __restrict__ g_left, const uint32_t* __restrict__ g_level,
                            const uint32_t* __restrict__ pre, uint32_t nn, uint32_t* ca, uint32_t* cb, uint32_t* pc,
                            uint32_t* lv, uint32_t* par, uint32_t* interior) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nn) return;
  const uint32_t i = pre[g];
  const uint32_t l = g_left[g];
  lv[i] = g_level[g];
  if (g == 0) par[i] = NONE;
  if (l == NONE) {
    ca[i] = g_start[g];
    cb[i] = 0;
    pc[i] = g_count[g];
    interior[i] = 0;
  } else {
    const uint32_t a = pre[l], b = pre[l + 1];
    ca[i] = a;
    cb[i] = b;
    pc[i] = 0;
    interior[i] = 1;
    par[a] = i;
    par[b] = i;
  }
}

// pack_tree's leaf ref: 0x80000000 | (count-1) << 27 | first (count <= 16).
__device__ __forceinline__ uint32_t leaf_ref_d(uint32_t first, uint32_t count) {
  return sqh::LEAF_BIT | ((count - 1) << 27) | first;
}

__device__ __forceinline__ uint32_t chain_len(uint32_t count) {
  return count > sqh::LEAF_MAX_COUNT ? (count + sqh::LEAF_MAX_COUNT - 1) / sqh::LEAF_MAX_COUNT - 1 : 0u;
}

// Synthetic pairs each interior node's leaf children need (pack_tree's
// leaf_chain), slot a then slot b.
__global__ void k_chain_counts(const uint32_t* __restrict__ ca, const uint32_t* __restrict__ cb, const uint32_t* __restrict__ pc,
                               const uint32_t* __restrict__ interior, const uint32_t* __restrict__ pair_of, uint32_t nn,
                               uint32_t* chains) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn || !interior[i]) return;
  chains[pair_of[i]] = chain_len(pc[ca[i]]) + chain_len(pc[cb[i]]);
}

__device__ __forceinline__ void put_ref(float* nd, int D, int slot, uint32_t ref) {
  nd[4 * D + slot] = __uint_as_float(ref);
}

// A leaf of `count` primitives at `first` under pair `parent` (depth d):
// a leaf ref, or a chain of synthetic pairs starting at pair q (pack_tree).
__device__ uint32_t write_chain(float* nodes, int D, uint32_t* depth, uint32_t first, uint32_t count, uint32_t parent,
                                uint32_t d, uint32_t q) {
  if (count <= sqh::LEAF_MAX_COUNT) return leaf_ref_d(first, count);
  const int NF = (D + 1) * 4;
  const uint32_t head = q;
  while (count > sqh::LEAF_MAX_COUNT) {
    float* nd = nodes + (size_t)q * NF;
    put_ref(nd, D, 0, leaf_ref_d(first, sqh::LEAF_MAX_COUNT));
    first += sqh::LEAF_MAX_COUNT;
    count -= sqh::LEAF_MAX_COUNT;
    put_ref(nd, D, 1, count > sqh::LEAF_MAX_COUNT ? q + 1 : leaf_ref_d(first, count));
    put_ref(nd, D, 2, parent);
    put_ref(nd, D, 3, 1u);
    depth[q] = d;
    parent = q;
    ++q;
    ++d;
  }
  return head;
}

// Pair nodes of the interior reference nodes (boxes are left to the refit).
__global__ void k_fill_pairs(const uint32_t* __restrict__ ca, const uint32_t* __restrict__ cb, const uint32_t* __restrict__ pc,
                             const uint32_t* __restrict__ lv, const uint32_t* __restrict__ par,
                             const uint32_t* __restrict__ interior, const uint32_t* __restrict__ pair_of,
                             const uint32_t* __restrict__ chain_base, uint32_t nn, uint32_t np, int D, float* nodes,
                             uint32_t* depth) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn || !interior[i]) return;
  const int NF = (D + 1) * 4;
  const uint32_t p = pair_of[i];
  float* nd = nodes + (size_t)p * NF;
  uint32_t q = np + chain_base[p];
  const uint32_t kids[2] = {ca[i], cb[i]};
  for (int s = 0; s < 2; ++s) {
    const uint32_t c = kids[s];
    uint32_t ref;
    if (interior[c]) {
      ref = pair_of[c];
    } else {
      ref = write_chain(nodes, D, depth, ca[c], pc[c], p, lv[i] + 1, q);
      q += chain_len(pc[c]);
    }
    put_ref(nd, D, s, ref);
  }
  put_ref(nd, D, 2, i == 0 ? NONE : pair_of[par[i]]);
  put_ref(nd, D, 3, 0u);
  depth[p] = lv[i];
}

// Root record: ref (bounds come from the refit).  A leaf root chains from pair 0.
__global__ void k_root_ref(const uint32_t* __restrict__ pc, uint32_t n_prims, int D, float* nodes, uint32_t* depth, float* root) {
  uint32_t ref = 0;
  if (pc[0] != 0) ref = write_chain(nodes, D, depth, 0, n_prims, NONE, 0, 0);
  root[2 * D] = __uint_as_float(ref);
}

// Deepest pair: a wave maximum, then one atomic per wave.
__global__ __launch_bounds__(256) void k_depth_max(const uint32_t* __restrict__ depth, uint32_t np, uint32_t* maxd) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t v = p < np ? depth[p] : 0u;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off));
  if ((threadIdx.x & 63u) == 0) atomicMax(maxd, v);
}

// Pairs per depth: counted in LDS per block (for up to HIST_LDS depths), then
// one global add per block and depth.
constexpr uint32_t HIST_LDS = 2048;
__global__ __launch_bounds__(256) void k_depth_hist(const uint32_t* __restrict__ depth, uint32_t np, uint32_t bins, uint32_t* hist) {
  __shared__ uint32_t h[HIST_LDS];
  const bool local = bins <= HIST_LDS;
  if (local)
    for (uint32_t k = threadIdx.x; k < bins; k += 256) h[k] = 0;
  __syncthreads();
  for (uint32_t p = blockIdx.x * 256 * 16 + threadIdx.x, e = 0; e < 16; ++e, p += 256) {
    if (p >= np) break;
    if (local) atomicAdd(h + depth[p], 1u);
    else atomicAdd(hist + depth[p], 1u);
  }
  __syncthreads();
  if (local)
    for (uint32_t k = threadIdx.x; k < bins; k += 256)
      if (h[k]) atomicAdd(hist + k, h[k]);
}

__global__ void k_level_keys(const uint32_t* __restrict__ depth, uint32_t np, uint32_t maxd, uint32_t* key, uint32_t* val) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  key[p] = maxd - depth[p];
  val[p] = p;
}

__global__ void k_fill_words(uint32_t* p, uint64_t n, uint32_t v) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

int bits_for(uint64_t v) {
  int b = 0;
  while (b < 32 && (1ull << b) <= v) ++b;
  return b;
}

}  // namespace

// Scratch of a device build: one DevBuf per array, kept in the context so that
// repeated rebuilds (moving scenes) allocate once.
enum BuildBuf {
  BB_KEYS, BB_IDX, BB_IDX2, BB_SEG, BB_NODE_OF, BB_K64, BB_K64B, BB_START, BB_COUNT, BB_LEFT, BB_LEVEL, BB_SIZE, BB_PRE,
  BB_SB, BB_FLAG, BB_AXIS, BB_RANK, BB_TMP, BB_INTERIOR, BB_PAIR_OF, BB_CHAINS, BB_CHAIN_BASE, BB_LV, BB_PAR, BB_DEPTH,
  BB_LKEY, BB_LKEY2, BB_SMALL, BB_COUNT_
};

// Exclusive scan of n u32 (scratch: BB_TMP, sized by the caller for the
// largest n), with the total copied back.
static int scan_u32(sq_ctx* c, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total_host) {
  DevBuf& tmp = c->build[BB_TMP];
  size_t bytes = 0;
  HIP_TRY(rocprim::exclusive_scan(nullptr, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), c->stream));
  if (bytes > tmp.bytes) return fail(SQ_ERR_HIP, "device build: scan scratch too small");
  HIP_TRY(rocprim::exclusive_scan(tmp.p, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), c->stream));
  if (total_host) {
    uint32_t last[2];
    HIP_TRY(hipMemcpyAsync(&last[0], out + n - 1, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(&last[1], in + n - 1, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *total_host = last[0] + last[1];
  }
  return SQ_OK;
}

// Builds the tree of the n device primitives into b (dim, kind, n_prims set;
// b->ctx the context): device arrays, host metadata, stats; the host
// reference tree is left empty (ensure_host_tree).
int device_build(sq_bvh* b, const float* prims) {
  sq_ctx* c = b->ctx;
  const int D = b->dim;
  const uint32_t n = (uint32_t)b->n_prims;
  hipStream_t st = c->stream;
  DevBuf* B = c->build;
  static_assert(BB_COUNT_ <= SQ_BUILD_BUFS, "build scratch");
  int rc;
  const uint64_t cap = 2ull * n + 2;        // nodes of the reference tree: fewer than 2n
  struct Need { int id; uint64_t bytes; } needs[] = {
      {BB_KEYS, (uint64_t)D * n * 4}, {BB_IDX, n * 4ull}, {BB_IDX2, n * 4ull}, {BB_SEG, n * 4ull}, {BB_NODE_OF, n * 4ull},
      {BB_K64, n * 8ull}, {BB_K64B, n * 8ull}, {BB_START, cap * 4}, {BB_COUNT, cap * 4}, {BB_LEFT, cap * 4},
      {BB_LEVEL, cap * 4}, {BB_SIZE, cap * 4}, {BB_PRE, cap * 4}, {BB_SB, cap * 2 * D * 4}, {BB_FLAG, cap * 4},
      {BB_AXIS, cap * 4}, {BB_RANK, cap * 4}, {BB_INTERIOR, cap * 4}, {BB_PAIR_OF, cap * 4}, {BB_CHAINS, cap * 4},
      {BB_CHAIN_BASE, cap * 4}, {BB_LV, cap * 4}, {BB_PAR, cap * 4}, {BB_DEPTH, cap * 4}, {BB_LKEY, cap * 4},
      {BB_LKEY2, cap * 4}, {BB_SMALL, 64 * 4}};
  for (const auto& nd : needs)
    if ((rc = B[nd.id].ensure(nd.bytes + 64))) return rc;
  if ((rc = b->ref_ca.ensure(cap * 4)) || (rc = b->ref_cb.ensure(cap * 4)) || (rc = b->ref_pc.ensure(cap * 4))) return rc;
  auto U = [&](int id) { return B[id].as<uint32_t>(); };
  uint32_t* keys = U(BB_KEYS);
  uint32_t* idx = U(BB_IDX);
  uint32_t* idx2 = U(BB_IDX2);
  uint32_t* seg = U(BB_SEG);
  uint32_t* node_of = U(BB_NODE_OF);
  uint32_t* g_start = U(BB_START);
  uint32_t* g_count = U(BB_COUNT);
  uint32_t* g_left = U(BB_LEFT);
  uint32_t* g_level = U(BB_LEVEL);

  rc = dispatch_dk(D, b->kind, [&]<int DD, int K>() -> int {
    k_centroid_keys<DD, K><<<grid_for(n, 256), 256, 0, st>>>(prims, n, keys);
    HIP_TRY(hipGetLastError());
    return SQ_OK;
  });
  if (rc) return rc;
  k_build_init<<<grid_for(n, 256), 256, 0, st>>>(n, idx, seg, node_of, g_start, g_count, g_level);
  HIP_TRY(hipGetLastError());

  // level-synchronous split
  std::vector<uint32_t> level_base{0};
  uint32_t base = 0, m = 1, level = 0;
  const int sort_bits = 32 + bits_for(n);
  size_t sort_tmp = 0, scan_tmp = 0;
  HIP_TRY(rocprim::radix_sort_pairs(nullptr, sort_tmp, B[BB_K64].as<uint64_t>(), B[BB_K64B].as<uint64_t>(), idx, idx2, n, 0,
                                    sort_bits, st));
  HIP_TRY(rocprim::exclusive_scan(nullptr, scan_tmp, idx, idx2, 0u, (size_t)cap, rocprim::plus<uint32_t>(), st));
  if ((rc = B[BB_TMP].ensure(std::max(sort_tmp, scan_tmp) + 64))) return rc;
  while (true) {
    uint32_t* start = g_start + base;
    uint32_t* count = g_count + base;
    uint32_t* flag = U(BB_FLAG);
    uint32_t* axis = U(BB_AXIS);
    uint32_t* rank = U(BB_RANK);
    uint32_t* sb = U(BB_SB);
    rc = dispatch_dk(D, KIND_AABB, [&]<int DD, int K>() -> int {
      k_seg_bounds_init<DD><<<grid_for(m, 256), 256, 0, st>>>(count, m, sb);
      k_seg_bounds_small<DD><<<grid_for(m, 4), 256, 0, st>>>(start, count, m, idx, keys, n, sb);
      // a median split leaves at most ceil(n / 2^level) primitives in a node
      if ((((uint64_t)n + (1ull << level) - 1) >> level) > SMALL_SEG)
        k_seg_bounds_large<DD><<<grid_for(n, TILE), 256, 0, st>>>(seg, node_of, count, idx, keys, n, sb);
      k_decide<DD><<<grid_for(m, 256), 256, 0, st>>>(count, m, sb, flag, axis);
      HIP_TRY(hipGetLastError());
      return SQ_OK;
    });
    if (rc) return rc;
    uint32_t splits = 0;
    if ((rc = scan_u32(c, flag, rank, m, &splits))) return rc;
    if (splits) {
      k_sort_keys<<<grid_for(n, 256), 256, 0, st>>>(seg, node_of, flag, axis, idx, keys, n, B[BB_K64].as<uint64_t>());
      HIP_TRY(hipGetLastError());
      HIP_TRY(rocprim::radix_sort_pairs(B[BB_TMP].p, sort_tmp, B[BB_K64].as<uint64_t>(), B[BB_K64B].as<uint64_t>(), idx, idx2,
                                        n, 0, sort_bits, st));
      std::swap(idx, idx2);
      k_split_positions<<<grid_for(n, 256), 256, 0, st>>>(seg, node_of, flag, count, n);
      HIP_TRY(hipGetLastError());
    }
    const uint32_t base_next = base + m;
    if (splits && (uint64_t)base_next + 2ull * splits > cap) return fail(SQ_ERR_HIP, "device build: node capacity exceeded");
    k_children<<<grid_for(m, 256), 256, 0, st>>>(start, count, m, base, level, flag, rank, g_left, g_start + base_next,
                                                 g_count + base_next, g_level + base_next, base_next, node_of);
    HIP_TRY(hipGetLastError());
    if (!splits) break;
    base = base_next;
    m = 2 * splits;
    ++level;
    level_base.push_back(base);
  }
  const uint32_t nn = base + m;                     // reference nodes
  const uint32_t levels = level + 1;
  level_base.push_back(nn);
  // the final order of the index array is the reference's prim_indices

  // pre-order numbering
  uint32_t* size = U(BB_SIZE);
  uint32_t* pre = U(BB_PRE);
  for (uint32_t l = levels; l-- > 0;) {
    const uint32_t lb = level_base[l], lm = level_base[l + 1] - lb;
    k_subtree_size<<<grid_for(lm, 256), 256, 0, st>>>(g_left, lb, lm, size);
  }
  for (uint32_t l = 0; l < levels; ++l) {
    const uint32_t lb = level_base[l], lm = level_base[l + 1] - lb;
    k_preorder<<<grid_for(lm, 256), 256, 0, st>>>(g_left, size, lb, lm, pre);
  }
  HIP_TRY(hipGetLastError());
  uint32_t* ca = b->ref_ca.as<uint32_t>();
  uint32_t* cb = b->ref_cb.as<uint32_t>();
  uint32_t* pc = b->ref_pc.as<uint32_t>();
  uint32_t* lv = U(BB_LV);
  uint32_t* par = U(BB_PAR);
  uint32_t* interior = U(BB_INTERIOR);
  k_ref_nodes<<<grid_for(nn, 256), 256, 0, st>>>(g_start, g_count, g_left, g_level, pre, nn, ca, cb, pc, lv, par, interior);
  HIP_TRY(hipGetLastError());

  // pair nodes
  uint32_t np = 0, n_synth = 0;
  uint32_t* pair_of = U(BB_PAIR_OF);
  uint32_t* chains = U(BB_CHAINS);
  uint32_t* chain_base = U(BB_CHAIN_BASE);
  if ((rc = scan_u32(c, interior, pair_of, nn, &np))) return rc;
  uint32_t root_pc = 0;
  HIP_TRY(hipMemcpyAsync(&root_pc, pc, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (np) {
    k_chain_counts<<<grid_for(nn, 256), 256, 0, st>>>(ca, cb, pc, interior, pair_of, nn, chains);
    HIP_TRY(hipGetLastError());
    if ((rc = scan_u32(c, chains, chain_base, np, &n_synth))) return rc;
  } else {
    n_synth = root_pc > sqh::LEAF_MAX_COUNT ? (root_pc + sqh::LEAF_MAX_COUNT - 1) / sqh::LEAF_MAX_COUNT - 1 : 0;
  }
  const uint32_t n_pairs = np + n_synth;
  const int NF = (D + 1) * 4;
  // room for any tree of n primitives (at most n - 1 interior pairs and n/16
  // chain pairs), so that a rebuild of the same count never moves the device
  // arrays that batch sets and captured graphs point at
  const uint64_t pair_cap = (uint64_t)n + n / sqh::LEAF_MAX_COUNT + 1;
  if (n_pairs > pair_cap) return fail(SQ_ERR_HIP, "device build: pair count out of range");
  if ((rc = b->nodes.ensure(pair_cap * NF * 4))) return rc;
  if ((rc = b->root.ensure(16 * 4))) return rc;
  HIP_TRY(hipMemsetAsync(b->nodes.p, 0, (size_t)std::max<uint32_t>(n_pairs, 1) * NF * 4, st));
  HIP_TRY(hipMemsetAsync(b->root.p, 0, 16 * 4, st));
  uint32_t* depth = U(BB_DEPTH);
  if (np) {
    k_fill_pairs<<<grid_for(nn, 256), 256, 0, st>>>(ca, cb, pc, lv, par, interior, pair_of, chain_base, nn, np, D,
                                                    b->nodes.as<float>(), depth);
    HIP_TRY(hipGetLastError());
  }
  k_root_ref<<<1, 1, 0, st>>>(pc, n, D, b->nodes.as<float>(), depth, b->root.as<float>());
  HIP_TRY(hipGetLastError());

  // depth levels, deepest first, pairs in index order within a level
  uint32_t maxd = 0;
  std::vector<uint32_t> hist;
  if (n_pairs) {
    // pair depths exceed the tree's levels only along leaf chains; the
    // histogram is sized by the deepest pair
    uint32_t* small = U(BB_SMALL);
    k_fill_words<<<1, 64, 0, st>>>(small, 1, 0u);
    k_depth_max<<<grid_for(n_pairs, 256), 256, 0, st>>>(depth, n_pairs, small);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(&maxd, small, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if ((rc = B[BB_LKEY2].ensure(((size_t)maxd + 1) * 4 + 64))) return rc;
    uint32_t* hist_d = U(BB_LKEY2);
    k_fill_words<<<grid_for(maxd + 1, 256), 256, 0, st>>>(hist_d, maxd + 1, 0u);
    k_depth_hist<<<grid_for(n_pairs, 256 * 16), 256, 0, st>>>(depth, n_pairs, maxd + 1, hist_d);
    HIP_TRY(hipGetLastError());
    hist.resize(maxd + 1);
    HIP_TRY(hipMemcpyAsync(hist.data(), hist_d, ((size_t)maxd + 1) * 4, hipMemcpyDeviceToHost, st));
    if ((rc = b->levels.ensure(pair_cap * 4))) return rc;
    uint32_t* lkey = U(BB_LKEY);
    uint32_t* lval = U(BB_CHAINS);                  // chains no longer needed
    k_level_keys<<<grid_for(n_pairs, 256), 256, 0, st>>>(depth, n_pairs, maxd, lkey, lval);
    HIP_TRY(hipGetLastError());
    size_t t2 = 0;
    uint32_t* lkey_out = U(BB_RANK);
    HIP_TRY(rocprim::radix_sort_pairs(nullptr, t2, lkey, lkey_out, lval, b->levels.as<uint32_t>(), n_pairs, 0,
                                      std::max(1, bits_for(maxd)), st));
    if (t2 + 64 > B[BB_TMP].bytes) {
      HIP_TRY(hipStreamSynchronize(st));
      if ((rc = B[BB_TMP].ensure(t2 + 64))) return rc;
    }
    HIP_TRY(rocprim::radix_sort_pairs(B[BB_TMP].p, t2, lkey, lkey_out, lval, b->levels.as<uint32_t>(), n_pairs, 0,
                                      std::max(1, bits_for(maxd)), st));
  }
  // ids: the index array
  if ((rc = b->ids.ensure((size_t)n * 4))) return rc;
  HIP_TRY(hipMemcpyAsync(b->ids.p, idx, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
  if ((rc = b->prims.ensure((size_t)n * sqh::device_prim_stride(D, b->kind) * 4))) return rc;
  HIP_TRY(hipStreamSynchronize(st));

  // host metadata (what queries and refits read)
  sqh::Packed& P = b->packed;
  P = sqh::Packed();
  P.dim = D;
  P.n_pairs = n_pairs;
  P.n_synth = n_synth;
  P.max_stack = n_pairs ? maxd + 2 : 1;
  P.level_off.push_back(0);
  for (size_t d = hist.size(); d-- > 0;) P.level_off.push_back(P.level_off.back() + hist[d]);
  b->ref_nodes = nn;
  b->tree = sqh::RefTree();
  b->tree.dim = D;
  b->host_tree_ok = false;
  b->stats = sq_tree_stats{};
  b->stats.node_count = nn;
  b->stats.leaf_count = nn - np;
  b->stats.prim_count = n;
  b->stats.max_depth = levels;
  b->stats.bytes = (uint64_t)nn * (8 * D + 12) + (uint64_t)n * 4;
  b->stats.quality = 0.0f;
  // primitives into leaf order and every box, as a refit does
  return sq_bvh_refit(b, prims);
}

// The host copy of the reference tree (and pack_tree's host arrays) of a
// device-built tree, downloaded on first use.
int ensure_host_tree(sq_bvh* b) {
  if (b->host_tree_ok) return SQ_OK;
  sq_ctx* c = b->ctx;
  const int D = b->dim;
  const size_t nn = b->ref_nodes, n = b->n_prims;
  sqh::RefTree t;
  t.dim = D;
  t.child_a.resize(nn);
  t.child_b.resize(nn);
  t.prim_count.resize(nn);
  t.prim_indices.resize(n);
  t.bounds.assign(nn * 2 * D, 0.0f);
  HIP_TRY(hipMemcpyAsync(t.child_a.data(), b->ref_ca.p, nn * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(t.child_b.data(), b->ref_cb.p, nn * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(t.prim_count.data(), b->ref_pc.p, nn * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(t.prim_indices.data(), b->ids.p, n * 4, hipMemcpyDeviceToHost, c->stream));
  const int NF = (D + 1) * 4;
  std::vector<float> nodes((size_t)b->packed.n_pairs * NF), root(16);
  if (!nodes.empty()) HIP_TRY(hipMemcpyAsync(nodes.data(), b->nodes.p, nodes.size() * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(root.data(), b->root.p, 64, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  sqh::Packed p;
  try {
    sqh::pack_tree(t, p);                                    // slot_of_ref from the topology
  } catch (const std::exception& e) {
    return fail(SQ_ERR_HIP, e.what());
  }
  // reference bounds = the device boxes (download_bounds' mapping)
  for (size_t i = 0; i < nn; ++i) {
    float* bo = t.bounds.data() + i * 2 * D;
    if (i == 0) {
      for (int k = 0; k < 2 * D; ++k) bo[k] = root[k];
      continue;
    }
    const uint32_t slot = p.slot_of_ref[i];
    const float* nd = nodes.data() + (size_t)(slot >> 1) * NF;
    const int s = (int)(slot & 1u);
    for (int k = 0; k < D; ++k) { bo[k] = nd[4 * k + 2 * s]; bo[D + k] = nd[4 * k + 2 * s + 1]; }
  }
  const uint32_t max_stack = b->packed.max_stack;
  const uint32_t n_synth = b->packed.n_synth;
  const auto level_off = b->packed.level_off;
  b->tree = std::move(t);
  try {
    sqh::pack_tree(b->tree, b->packed);
  } catch (const std::exception& e) {
    return fail(SQ_ERR_HIP, e.what());
  }
  b->packed.max_stack = max_stack;
  b->packed.n_synth = n_synth;
  b->packed.level_off = level_off;
  b->stats.quality = sqh::tree_interior_surface_sum(b->tree);
  b->host_tree_ok = true;
  return SQ_OK;
}

extern "C" {

int sq_bvh_build_device(sq_ctx* ctx, int dim, int kind, const float* prims_device, uint64_t n, sq_bvh** out) {
  if (!ctx || !out || (n && !prims_device)) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = nullptr;
  int rc = dispatch_dk(dim, kind, [&]<int D, int K>() { return SQ_OK; });
  if (rc) return rc;
  if (n > sqh::LEAF_FIRST_MASK) return fail(SQ_ERR_INVALID, "at most 2^27 primitives per tree");
  if (n == 0) return sq_bvh_build(ctx, dim, kind, nullptr, 0, out);
  HIP_TRY(hipSetDevice(ctx->device));
  auto b = std::make_unique<sq_bvh>();
  b->dim = dim;
  b->kind = kind;
  b->n_prims = n;
  b->ctx = ctx;
  if ((rc = device_build(b.get(), prims_device))) return rc;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx_retain(ctx);
  *out = b.release();
  return SQ_OK;
}

int sq_bvh_rebuild(sq_bvh* b, const float* prims_device) {
  if (!b || (b->n_prims && !prims_device)) return fail(SQ_ERR_INVALID, "NULL argument");
  if (!b->n_prims) return SQ_OK;
  int rc = device_build(b, prims_device);
  if (rc) return rc;
  b->auto_baseline = 0.0;          // sq_bvh_refit_auto takes the rebuilt tree as its new baseline
  HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  return SQ_OK;
}

}  // extern "C"
