// order.hip -- Morton ray order: rays sorted by a 30-bit key that interleaves
// the quantized components of origin and direction, each quantized within the
// batch's own bounds (found by a min/max pass), with the key's bits shared
// among the components that vary in the batch: a batch of camera rays (one
// origin) is ordered by direction alone, 10 bits per component, so that 64
// consecutive rays cover a small tile of the image; an incoherent batch gets 5
// bits per component of origin and direction.  Neighbouring lanes of a wave
// then walk similar paths (and a packet walk, packet.hip, shares one walk among
// them).  Only the visiting order changes: every ray's result is computed and
// written as before.
#include <rocprim/device/device_radix_sort.hpp>

#include "engine_internal.h"

namespace {

// Floats as unsigned integers in the same order (for atomic min / max).
__device__ __forceinline__ uint32_t ordered_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float from_ordered(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// Per-component minima and maxima of the batch (ordered bits): each block
// reduces its share into part[block][0..2D) (minima) and [2D..4D) (maxima),
// then one block reduces the parts into bounds[4D].  No atomics: the same
// bounds on every call.
constexpr int BOUND_BLOCKS = 256;

template <int D>
__global__ __launch_bounds__(256) void k_bounds_part(const float* __restrict__ q, uint32_t n, uint32_t qs, uint32_t* part) {
  constexpr int C = 2 * D;
  __shared__ uint32_t red[4][2 * C];
  uint32_t mn[C], mx[C];
#pragma unroll
  for (int k = 0; k < C; ++k) { mn[k] = 0xFFFFFFFFu; mx[k] = 0u; }
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float* r = q + (size_t)i * qs;
#pragma unroll
    for (int k = 0; k < C; ++k) {
      const float v = r[k];
      if (v == v) { const uint32_t o = ordered_bits(v); mn[k] = min(mn[k], o); mx[k] = max(mx[k], o); }
    }
  }
  const uint32_t w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < C; ++k) {
    for (int o = 32; o > 0; o >>= 1) {
      mn[k] = min(mn[k], (uint32_t)__shfl_xor((int)mn[k], o));
      mx[k] = max(mx[k], (uint32_t)__shfl_xor((int)mx[k], o));
    }
    if ((threadIdx.x & 63u) == 0) { red[w][k] = mn[k]; red[w][C + k] = mx[k]; }
  }
  __syncthreads();
  if (threadIdx.x < 2 * C) {
    const int k = threadIdx.x;
    uint32_t v = red[0][k];
    for (uint32_t j = 1; j < blockDim.x / 64; ++j) v = k < C ? min(v, red[j][k]) : max(v, red[j][k]);
    part[(size_t)blockIdx.x * 2 * C + k] = v;
  }
}

// Block k reduces component slot k (a minimum for k < comps, else a maximum)
// over the blocks' parts.
__global__ __launch_bounds__(256) void k_bounds_final(const uint32_t* __restrict__ part, uint32_t blocks, int comps,
                                                      uint32_t* bounds) {
  __shared__ uint32_t red[4];
  const int k = blockIdx.x;
  const bool is_min = k < comps;
  uint32_t v = is_min ? 0xFFFFFFFFu : 0u;
  for (uint32_t j = threadIdx.x; j < blocks; j += blockDim.x) {
    const uint32_t x = part[(size_t)j * 2 * comps + k];
    v = is_min ? min(v, x) : max(v, x);
  }
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t x = (uint32_t)__shfl_xor((int)v, o);
    v = is_min ? min(v, x) : max(v, x);
  }
  if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) v = is_min ? min(v, red[w]) : max(v, red[w]);
    bounds[k] = v;
  }
}

// key: the components with a positive, finite extent in the batch share 30
// bits (30 / count each), interleaved from the most significant bit level
// down, each quantized within its batch range.
template <int D>
__global__ void k_ray_keys(const float* __restrict__ q, uint32_t n, uint32_t qs, const uint32_t* __restrict__ bounds,
                           uint32_t* keys, uint32_t* idx) {
  constexpr int C = 2 * D;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float lo[C], inv[C];
  int active = 0;
#pragma unroll
  for (int k = 0; k < C; ++k) {
    lo[k] = from_ordered(bounds[k]);
    const float ext = from_ordered(bounds[C + k]) - lo[k];
    const bool on = ext > 0.0f && ext < SQ_INF;
    inv[k] = on ? 1.0f / ext : 0.0f;
    active += on ? 1 : 0;
  }
  const int bits = active ? 30 / active : 0;
  const float qmax = (float)((1u << bits) - 1u);
  const float* r = q + (size_t)i * qs;
  uint32_t c[C];
#pragma unroll
  for (int k = 0; k < C; ++k) {
    const float u = (r[k] - lo[k]) * inv[k];
    c[k] = inv[k] > 0.0f ? (uint32_t)(fminf(fmaxf(u, 0.0f), 1.0f) * qmax) : 0u;
  }
  uint32_t key = 0;
  for (int b = bits - 1; b >= 0; --b)
#pragma unroll
    for (int k = 0; k < C; ++k)
      if (inv[k] > 0.0f) key = (key << 1) | ((c[k] >> b) & 1u);
  keys[i] = key;
  idx[i] = i;
}

}  // namespace

int ray_order(sq_ctx* c, int dim, const float* q, uint32_t n, uint32_t qs, const uint32_t** order) {
  int rc;
  if ((rc = c->okeys.ensure((size_t)n * 4 + 4 * 2 * 6 * 4)) || (rc = c->okeys2.ensure((size_t)n * 4)) || (rc = c->oidx.ensure((size_t)n * 4)) ||
      (rc = c->oorder.ensure((size_t)n * 4)))
    return rc;
  uint32_t* bounds = c->okeys.as<uint32_t>() + n;   // 4D words after the keys
  const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>(BOUND_BLOCKS, (n + 255) / 256));
  if ((rc = c->oidx.ensure(std::max<size_t>((size_t)n * 4, (size_t)blocks * 4 * 6 * 4)))) return rc;
  switch (dim) {
#define SQ_KEYS(DD)                                                                                                   \
  case DD:                                                                                                            \
    k_bounds_part<DD><<<blocks, 256, 0, c->stream>>>(q, n, qs, c->oidx.as<uint32_t>());                              \
    k_bounds_final<<<4 * DD, 256, 0, c->stream>>>(c->oidx.as<uint32_t>(), blocks, 2 * DD, bounds);                   \
    k_ray_keys<DD><<<grid_for(n, 256), 256, 0, c->stream>>>(q, n, qs, bounds, c->okeys.as<uint32_t>(), c->oidx.as<uint32_t>()); \
    break;
    SQ_KEYS(2) SQ_KEYS(3) SQ_KEYS(4) SQ_KEYS(5) SQ_KEYS(6)
#undef SQ_KEYS
    default: return fail(SQ_ERR_UNSUPPORTED, "ray order: dimension not instantiated");
  }
  HIP_TRY(hipGetLastError());
  size_t temp = 0;
  HIP_TRY(rocprim::radix_sort_pairs(nullptr, temp, c->okeys.as<uint32_t>(), c->okeys2.as<uint32_t>(), c->oidx.as<uint32_t>(),
                                    c->oorder.as<uint32_t>(), n, 0, 30, c->stream));
  if ((rc = c->otemp.ensure(temp + 16))) return rc;
  HIP_TRY(rocprim::radix_sort_pairs(c->otemp.p, temp, c->okeys.as<uint32_t>(), c->okeys2.as<uint32_t>(), c->oidx.as<uint32_t>(),
                                    c->oorder.as<uint32_t>(), n, 0, 30, c->stream));
  *order = c->oorder.as<uint32_t>();
  return SQ_OK;
}
