// order.hip -- Morton ray order: rays sorted by a 30-bit key that interleaves
// the quantized origin (in the scene box) with the quantized direction, so
// that neighbouring lanes of a wave walk similar paths.  Only the visiting
// order changes: every ray's result is computed and written as before.
#include <rocprim/device/device_radix_sort.hpp>

#include "engine_internal.h"

namespace {

// key: BITS bits per component, components o[0..D), d[0..D), interleaved
// from the most significant bit level down.  Origins are clamped to the box.
template <int D>
__global__ void k_ray_keys(const float* __restrict__ q, uint32_t n, uint32_t qs, const float* __restrict__ box, uint32_t* keys,
                           uint32_t* idx) {
  constexpr int BITS = 30 / (2 * D);
  constexpr float QMAX = (float)((1 << BITS) - 1);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* r = q + (size_t)i * qs;
  uint32_t c[2 * D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const float ext = box[D + k] - box[k];
    const float u = ext > 0.0f ? (r[k] - box[k]) / ext : 0.0f;
    c[k] = (uint32_t)(fminf(fmaxf(u, 0.0f), 1.0f) * QMAX);
    const float v = (r[D + k] + 1.0f) * 0.5f;
    c[D + k] = (uint32_t)(fminf(fmaxf(v, 0.0f), 1.0f) * QMAX);
  }
  uint32_t key = 0;
#pragma unroll
  for (int b = BITS - 1; b >= 0; --b)
#pragma unroll
    for (int k = 0; k < 2 * D; ++k) key = (key << 1) | ((c[k] >> b) & 1u);
  keys[i] = key;
  idx[i] = i;
}

}  // namespace

int ray_order(sq_ctx* c, int dim, const float* q, uint32_t n, uint32_t qs, const float* box, const uint32_t** order) {
  int rc;
  if ((rc = c->okeys.ensure((size_t)n * 4)) || (rc = c->okeys2.ensure((size_t)n * 4)) || (rc = c->oidx.ensure((size_t)n * 4)) ||
      (rc = c->oorder.ensure((size_t)n * 4)))
    return rc;
  switch (dim) {
#define SQ_KEYS(DD) case DD: k_ray_keys<DD><<<grid_for(n, 256), 256, 0, c->stream>>>(q, n, qs, box, c->okeys.as<uint32_t>(), c->oidx.as<uint32_t>()); break;
    SQ_KEYS(2) SQ_KEYS(3) SQ_KEYS(4) SQ_KEYS(5) SQ_KEYS(6)
#undef SQ_KEYS
    default: return fail(SQ_ERR_UNSUPPORTED, "ray order: dimension not instantiated");
  }
  HIP_TRY(hipGetLastError());
  const int bits = 2 * dim * (30 / (2 * dim));
  size_t temp = 0;
  HIP_TRY(rocprim::radix_sort_pairs(nullptr, temp, c->okeys.as<uint32_t>(), c->okeys2.as<uint32_t>(), c->oidx.as<uint32_t>(),
                                    c->oorder.as<uint32_t>(), n, 0, bits, c->stream));
  if ((rc = c->otemp.ensure(temp + 16))) return rc;
  HIP_TRY(rocprim::radix_sort_pairs(c->otemp.p, temp, c->okeys.as<uint32_t>(), c->okeys2.as<uint32_t>(), c->oidx.as<uint32_t>(),
                                    c->oorder.as<uint32_t>(), n, 0, bits, c->stream));
  *order = c->oorder.as<uint32_t>();
  return SQ_OK;
}
