// packet.hip -- wave-coherent (packet) nearest-hit walk for batches of
// coherent rays (camera rays: one origin, neighbouring directions).
//
// One wave walks the tree once for its 64 rays: the node and the triangles
// under test are the same for every lane, so they arrive by scalar loads
// (one 64-byte s_load per pair node instead of 64 vector fetches) and every
// lane tests them against its own ray with its own limit.  A lane takes part
// in a node only when the node passes that lane's own test, so each lane
// visits exactly the nodes its own walk visits, in the same order:
//   * exact (SQ_HITS_REFERENCE, leaves whose hits may leave their box): the
//     reference's order (bvh.rs:73-124), child b's subtree first, child a
//     pushed with each lane's entry and taken at its pop by the lanes whose
//     entry is within their current best (the pop-time test, see order_pair);
//     the pieces of an over-full leaf in leaf order, untested;
//   * otherwise (guarded kinds, AABB leaves) any order gives each lane's
//     answer, and the packet descends first into the child more of its lanes
//     enter first.
// The wave's stack holds the node refs once (LDS) and one entry per lane
// (LDS, NaN for a lane that did not pass); a lane that finished its ray idles
// until the wave's last ray is done.  Incoherent batches walk better one ray
// per lane (pool.h): the caller chooses.
#include "engine_internal.h"

namespace {

typedef const float __attribute__((address_space(4)))* const_fp;

// N consecutive floats at a wave-uniform address, by scalar loads.
template <int N>
__device__ __forceinline__ void uniform_load(const float* p, float* out) {
  const_fp c = (const_fp)p;
#pragma unroll
  for (int k = 0; k < N; ++k) out[k] = c[k];
}

template <int D, int KIND, bool REFO>
__global__ __launch_bounds__(256) void k_nearest_packet(SceneView s, const float* __restrict__ q, uint32_t n, float max_toi,
                                                        const float* __restrict__ tmax, NearestOut out, uint32_t S,
                                                        const uint32_t* __restrict__ order) {
  constexpr int NF = (D + 1) * 4;
  constexpr int PS = Prim<D, KIND>::STRIDE;
  extern __shared__ uint32_t lds[];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t* sref = lds + (size_t)wave * S * 65;                 // [S] node refs
  float* sent = reinterpret_cast<float*>(sref + S);             // [S][64] lane entries
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = slot < n;
  const uint32_t i = valid && order ? order[slot] : slot;   // the ray this lane walks
  RayPolicy<D, KIND, !REFO && KIND != KIND_AABB ? true : false> pol;
  float limit = -1.0f;
  bool act = false;
  if (valid) {
    pol.setup(q + (size_t)i * 2 * D);
    limit = tmax ? tmax[i] : max_toi;
    act = pol.root(s.root, limit);
  }
  NearestResult res{SQ_INF, NO_LEAF, NO_LEAF};
  uint32_t cur = fbits(s.root[2 * D]);
  uint32_t sp = 0;
  while (true) {
    if (__ballot(act)) {
      cur = __builtin_amdgcn_readfirstlane(cur);
      if (!(cur & LEAF_BIT)) {
        float nd[NF];
        uniform_load<NF>(s.nodes + (size_t)cur * NF, nd);
        const uint32_t ra = fbits(nd[4 * D + 0]), rb = fbits(nd[4 * D + 1]);
        bool ha = false, hb = false;
        float ta = SQ_INF, tb = SQ_INF;
        uint32_t first = ra, second = rb;
        bool hf, hs;
        float ts;
        if (synthetic_pair<D>(nd)) {
          // pieces of one reference leaf: every lane at it tests them all, in leaf order
          hf = hs = act;
          ts = -SQ_INF;
        } else {
          if (act) pol.pair(nd, limit, ha, hb, ta, tb);
          bool a_first = false;
          if constexpr (!REFO) {
            const uint32_t pa = (uint32_t)__popcll(__ballot(ha && (!hb || ta <= tb)));
            const uint32_t pb = (uint32_t)__popcll(__ballot(hb && (!ha || tb < ta)));
            a_first = pa >= pb;
          }
          first = a_first ? ra : rb; hf = a_first ? ha : hb;
          second = a_first ? rb : ra; hs = a_first ? hb : ha; ts = a_first ? tb : ta;
        }
        const bool any_f = __ballot(hf) != 0, any_s = __ballot(hs) != 0;
        if (any_f && any_s) {
          sref[sp] = second;
          sent[sp * 64 + lane] = hs ? ts : __builtin_nanf("");
          ++sp;
          cur = first;
          act = hf;
          continue;
        }
        if (any_f) { cur = first; act = hf; continue; }
        if (any_s) { cur = second; act = hs; continue; }
      } else {
        const uint32_t first = cur & FIRST_MASK;
        const uint32_t cnt = ((cur >> 27) & 15u) + 1u;
        for (uint32_t j = 0; j < cnt; ++j) {
          const uint32_t pos = first + j;
          float pr[PS];
          uniform_load<PS>(s.prims + (size_t)pos * PS, pr);
          float t;
          if (act && pol.narrow(pr, limit, t)) {
            const uint32_t leaf = s.ids[pos];
            if (res.leaf == NO_LEAF || hit_before(t, leaf, res.t, res.leaf)) {
              res.t = t; res.leaf = leaf; res.pos = pos; limit = t;
            }
          }
        }
      }
    }
    // pop the next node some lane still takes
    bool got = false;
    while (sp > 0) {
      --sp;
      const float e = sent[sp * 64 + lane];
      act = valid && e <= limit;
      if (__ballot(act)) { cur = sref[sp]; got = true; break; }
    }
    if (!got) break;
  }
  if (!valid) return;
  const bool hit = res.leaf != NO_LEAF;
  out.leaf[i] = hit ? (int32_t)res.leaf : -1;
  out.toi[i] = res.t;
  if (out.point) {
#pragma unroll
    for (int k = 0; k < D; ++k) out.point[(size_t)i * D + k] = hit ? pol.r.o[k] + pol.r.d[k] * res.t : 0.0f;
  }
  if (out.normal) {
    float nv[D];
    if (hit) hit_normal<D, KIND>(pol.r, s.prims + (size_t)res.pos * PS, res.t, tmax ? tmax[i] : max_toi, nv);
    else for (int k = 0; k < D; ++k) nv[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < D; ++k) out.normal[(size_t)i * D + k] = nv[k];
  }
  if (out.sub) out.sub[i] = (hit && KIND != KIND_AABB) ? (int32_t)res.leaf : -1;
}

}  // namespace

// Nearest hits of rays q[n][2D] by the packet walk (no filter); exact: the
// reference's hits and order (else guarded hits, any order).  With `order`,
// the 64 rays of a wave are order[64w .. 64w + 63].
int packet_nearest(sq_bvh* b, const float* q, uint64_t n, float max_toi, const float* tmax, NearestOut out, bool exact,
                   const uint32_t* order) {
  sq_ctx* c = b->ctx;
  const uint32_t S = b->packed.max_stack;
  const size_t lds = (size_t)4 * S * 65 * 4;
  if (lds > 64 * 1024) return fail(SQ_ERR_INVALID, "packet walk: traversal stack too deep for LDS");
  const uint32_t blocks = (uint32_t)((n + 255) / 256);
  return dispatch_dk(b->dim, b->kind, [&]<int D, int K>() -> int {
    if (exact && K != KIND_AABB)
      k_nearest_packet<D, K, true><<<blocks, 256, lds, c->stream>>>(view_of(b), q, (uint32_t)n, max_toi, tmax, out, S, order);
    else
      k_nearest_packet<D, K, false><<<blocks, 256, lds, c->stream>>>(view_of(b), q, (uint32_t)n, max_toi, tmax, out, S, order);
    HIP_TRY(hipGetLastError());
    return SQ_OK;
  });
}
