// tlas.hip -- the two-level tree (crates/spatial-query/src/accel/tlas.rs) on
// the device: a top-level pair-node tree over instance world AABBs, instance
// records (rigid transform + kind data) and, for mesh instances, one device
// BVH per mesh walked in the instance's frame.
#include <cmath>
#include <cstring>

#include "engine_internal.h"
#include "pool.h"

namespace {

struct Inst {
  int dim, kind;
  uint64_t n;
  std::vector<float> rot, trans, halves, scale;   // by instance id (halves: radii for BALL)
  std::vector<uint32_t> mesh_of;
  std::vector<sq_bvh*> meshes;                    // borrowed
};

uint32_t record_stride(int dim, int kind) {
  uint32_t f = kind == INST_BOX ? dim * dim + 2 * dim : (kind == INST_BALL ? dim * dim + dim + 1 : dim * dim + dim + 2);
  return (f + 3) & ~3u;
}

// Isometry::transform_aabb(local_aabb) (isometry.rs:84-97) on the host, with
// the reference's operation order.
void world_boxes(const Inst& in, std::vector<float>& out) {
  const int D = in.dim;
  out.resize(in.n * 2 * D);
  for (uint64_t i = 0; i < in.n; ++i) {
    float lmin[sqh::MAX_DIM], lmax[sqh::MAX_DIM];
    if (in.kind == INST_BOX) {
      for (int k = 0; k < D; ++k) { lmax[k] = in.halves[i * D + k]; lmin[k] = lmax[k] * -1.0f; }
    } else if (in.kind == INST_BALL) {
      // local_aabb of a ball (shape_overlap.rs:320-322)
      for (int k = 0; k < D; ++k) { lmax[k] = in.halves[i]; lmin[k] = -in.halves[i]; }
    } else {
      // local_aabb of a scaled mesh: the mesh root box times the scale
      const sq_bvh* m = in.meshes[in.mesh_of[i]];
      const float s = in.scale[i];
      for (int k = 0; k < D; ++k) { lmin[k] = m->tree.bounds[k] * s; lmax[k] = m->tree.bounds[D + k] * s; }
    }
    const float* R = in.rot.data() + i * D * D;
    const float* t = in.trans.data() + i * D;
    float c[sqh::MAX_DIM], h[sqh::MAX_DIM], rc[sqh::MAX_DIM];
    for (int k = 0; k < D; ++k) { c[k] = (lmin[k] + lmax[k]) * 0.5f; h[k] = (lmax[k] - lmin[k]) * 0.5f; }
    for (int a = 0; a < D; ++a) {
      float acc = 0.0f;
      for (int b = 0; b < D; ++b) acc = acc + R[a * D + b] * c[b];
      rc[a] = acc + t[a];
    }
    float* o = out.data() + i * 2 * D;
    for (int a = 0; a < D; ++a) {
      float acc = 0.0f;
      for (int b = 0; b < D; ++b) acc = acc + fabsf(R[a * D + b]) * h[b];
      o[a] = rc[a] - acc;
      o[D + a] = rc[a] + acc;
    }
  }
}

void records(const Inst& in, std::vector<float>& out) {
  const int D = in.dim;
  const uint32_t xs = record_stride(D, in.kind);
  out.assign(in.n * xs, 0.0f);
  for (uint64_t i = 0; i < in.n; ++i) {
    float* x = out.data() + i * xs;
    std::memcpy(x, in.rot.data() + i * D * D, D * D * 4);
    std::memcpy(x + D * D, in.trans.data() + i * D, D * 4);
    if (in.kind == INST_BOX) {
      std::memcpy(x + D * D + D, in.halves.data() + i * D, D * 4);
    } else if (in.kind == INST_BALL) {
      x[D * D + D] = in.halves[i];
    } else {
      x[D * D + D] = 1.0f / in.scale[i];
      std::memcpy(&x[D * D + D + 1], &in.mesh_of[i], 4);
    }
  }
}

}  // namespace

struct sq_tlas {
  sq_ctx* ctx = nullptr;
  Inst in;
  sq_bvh* top = nullptr;                  // AABB-kind tree over world boxes
  DevBuf xf, world, blas;
  float baseline = 0.0f;
  uint32_t blas_stack = 0;
  ~sq_tlas() {
    if (top) sq_bvh_destroy(top);
  }
};

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

template <int D, int IK, bool ANY, bool EXACT>
__global__ __launch_bounds__(1024) void k_tlas_nearest(TlasView tv, const float* __restrict__ q, uint32_t n, float max_toi,
                                                       const float* __restrict__ tmax, NearestOut out, uint32_t s_top,
                                                       uint32_t s_bot) {
  extern __shared__ uint32_t lds[];
  const uint32_t B = blockDim.x, tid = threadIdx.x;
  const uint32_t i = blockIdx.x * B + tid;
  if (i >= n) return;
  RayD<D> r;
  ray_setup<D>(r, q + (size_t)i * 2 * D);
  const float lim = tmax ? tmax[i] : max_toi;
  const uint32_t S = s_top + s_bot;
  uint32_t* refs = lds + tid;
  float* ts = reinterpret_cast<float*>(lds + S * B) + tid;
  Stack st{refs, ts, B};
  Stack sb{refs + s_top * B, ts + s_top * B, B};
  TlasHit h = walk_tlas<D, IK, ANY, EXACT>(tv, r, lim, st, sb);
  const bool hit = h.inst != NO_LEAF;
  if (ANY) { out.any[i] = hit ? 1 : 0; return; }
  out.leaf[i] = hit ? (int32_t)h.inst : -1;
  out.toi[i] = h.t;
  if (out.point) {
#pragma unroll
    for (int k = 0; k < D; ++k) out.point[(size_t)i * D + k] = hit ? r.o[k] + r.d[k] * h.t : 0.0f;
  }
  if (out.normal) {
    float nv[D];
#pragma unroll
    for (int k = 0; k < D; ++k) nv[k] = 0.0f;
    if (hit) tlas_normal<D, IK>(tv, r, h, lim, nv);
#pragma unroll
    for (int k = 0; k < D; ++k) out.normal[(size_t)i * D + k] = nv[k];
  }
  if (out.sub) out.sub[i] = (hit && h.sub != NO_LEAF) ? (int32_t)h.sub : -1;
}

// Pooled, persistent form of k_tlas_nearest (pool_walk_tlas): waves take
// chunks of `chunk` rays from counter[0] (zero at the start, left zero by
// pool_retire), both stacks S-entry LDS rings spilling to their own columns.
// EXACT: the reference's hits (tlas_step); else the guarded walk.
template <int D, int IK, bool ANY, bool EXACT>
__global__ __launch_bounds__(256) void k_tlas_nearest_pool(TlasView tv, const float* __restrict__ q, uint32_t n, float max_toi,
                                                           const float* __restrict__ tmax, NearestOut out, uint32_t chunk,
                                                           uint32_t* counter, uint32_t S, SpillEntry* spill_top,
                                                           SpillEntry* spill_bot, uint32_t spill_stride, uint32_t pause_below) {
  extern __shared__ uint32_t lds[];
  const uint32_t B = blockDim.x, tid = threadIdx.x;
  const uint32_t gt = blockIdx.x * B + tid;
  RingStack st{lds + tid, reinterpret_cast<float*>(lds + S * B) + tid, B, S, spill_top + gt, spill_stride};
  RingStack sb{lds + 2 * S * B + tid, reinterpret_cast<float*>(lds + 3 * S * B) + tid, B, S, spill_bot + gt, spill_stride};
  const uint64_t waves = (uint64_t)gridDim.x * (B >> 6);
  const uint64_t home = (uint64_t)blockIdx.x * (B >> 6) + (tid >> 6);
  pool_walk_tlas<D, IK, ANY, Chunks, EXACT>(tv, q, Chunks{counter, n, chunk, home, waves}, max_toi, tmax, out, st, sb, pause_below);
  pool_retire(counter);
}

// Proposal-seeded two-level nearest walk (D = 3) over guarded hits
// (SQ_HITS_GUARDED): the proposed top-tree subtrees top[ray][0..3] (ids >=
// n_sub mark empty slots) are walked first, each from its own ref, with the
// shrinking limit; then the full walk runs from the root with that limit and
// best hit, skipping the subtrees already walked.  A guarded hit lies inside
// its instance's world box (and, for meshes, its triangle's box), so the
// answer does not depend on which subtrees go first: it is the guarded
// k_tlas_nearest_pool's.  top == NULL walks from the root at once.  COUNT adds
// to stats[0] the rays whose proposed subtrees held a hit.
template <int IK, bool COUNT>
__global__ __launch_bounds__(256) void k_tlas_nearest_proposed(TlasView tv, const float* __restrict__ q, uint32_t n, float max_toi,
                                                               const uint16_t* __restrict__ top, const uint32_t* __restrict__ sub_ref,
                                                               uint32_t n_sub, int32_t* out_inst, float* out_toi, int32_t* out_sub,
                                                               uint32_t chunk, uint32_t* counter, uint32_t S, SpillEntry* spill_top,
                                                               SpillEntry* spill_bot, uint32_t spill_stride, uint32_t pause_below,
                                                               unsigned long long* stats) {
  constexpr int D = 3;
  extern __shared__ uint32_t lds[];
  const uint32_t B = blockDim.x, tid = threadIdx.x;
  const uint32_t gt = blockIdx.x * B + tid;
  RingStack st{lds + tid, reinterpret_cast<float*>(lds + S * B) + tid, B, S, spill_top + gt, spill_stride};
  RingStack sb{lds + 2 * S * B + tid, reinterpret_cast<float*>(lds + 3 * S * B) + tid, B, S, spill_bot + gt, spill_stride};
  const uint64_t waves = (uint64_t)gridDim.x * (B >> 6);
  const uint64_t home = (uint64_t)blockIdx.x * (B >> 6) + (tid >> 6);
  Chunks src{counter, n, chunk, home, waves};
  const uint32_t lane = tid & 63u;
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  const uint32_t root_ref = fbits(tv.top.root[2 * D]);
  RayD<D> r;
  TlasWalk<D, false> w;
  uint32_t rid = NO_LEAF, seeded = 0;
  uint32_t done[4];   // subtrees this ray has walked; the root ref (never a child) fills empty slots
  int phase = 0;      // 0..3: a proposed subtree, 4: the full walk from the root
  uint64_t next = 0, end = 0;
  bool exhausted = !src.get(next, end);
  auto start = [&](int ph) {
    w.sp = 0;
    w.phase = 0;
    if (ph == 0)
#pragma unroll
      for (int k = 0; k < 4; ++k) done[k] = root_ref;
    if (top)
      for (; ph < 4; ++ph) {
        const uint32_t c = top[(size_t)rid * 4 + ph];
        if (c < n_sub) {
          phase = ph;
          w.cur = sub_ref[c];
#pragma unroll
          for (int k = 0; k < 4; ++k) done[k] = k == ph ? w.cur : done[k];
          return;
        }
      }
    phase = 4;
    if (w.res.inst != NO_LEAF) ++seeded;
    w.cur = root_ref;
  };
  auto write = [&](uint32_t i) {
    const bool hit = w.res.inst != NO_LEAF;
    out_inst[i] = hit ? (int32_t)w.res.inst : -1;
    out_toi[i] = w.res.t;
    if (out_sub) out_sub[i] = (hit && w.res.sub != NO_LEAF) ? (int32_t)w.res.sub : -1;
  };
  while (true) {
    bool need = rid == NO_LEAF && !exhausted;
    uint64_t m = __ballot(need);
    while (m) {
      if (next >= end) {
        if (!src.get(next, end)) { exhausted = true; break; }
        continue;
      }
      const uint64_t avail = end - next;
      const uint64_t rank = (uint64_t)__popcll(m & lt_mask);
      if (need && rank < avail) {
        need = false;
        rid = (uint32_t)(next + rank);
        ray_setup<D>(r, q + (size_t)rid * 2 * D);
        if (!tlas_begin<D>(tv, r, max_toi, w)) {
          write(rid);
          rid = NO_LEAF;
        } else {
          start(0);
        }
      }
      const uint64_t taken = (uint64_t)__popcll(m);
      next += taken < avail ? taken : avail;
      m = __ballot(need);
    }
    if (__ballot(!(exhausted && rid == NO_LEAF)) == 0) break;
    const uint32_t active = (uint32_t)__popcll(__ballot(rid != NO_LEAF));
    if (rid == NO_LEAF) continue;
    if (!tlas_step<D, IK, false, false, true>(tv, r, st, sb, w, WavePause{exhausted ? 0u : pause_below, active}, done)) continue;
    if (phase < 4) { start(phase + 1); continue; }   // this subtree is exhausted
    write(rid);
    rid = NO_LEAF;
  }
  if constexpr (COUNT) {
    unsigned long long v = seeded;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0 && v) atomicAdd(stats, v);
  }
  pool_retire(counter);
}

// MODE 0: raycast_all (one hit per instance, tlas.rs:158-215); MODE 2: overlap
// (every instance of a touched leaf node passing its test_overlap: exact for
// balls, the default true for the other kinds, tlas.rs:393-432); MODE 3:
// shapecast_all (tlas.rs:353-391, casts [n][4D]).
template <int D, int IK, int MODE, bool FILL>
__global__ __launch_bounds__(1024) void k_tlas_collect(TlasView tv, const float* __restrict__ q, uint32_t n, float max_toi,
                                                       uint32_t* counts, const uint64_t* __restrict__ offsets, uint64_t* keys,
                                                       uint32_t s_top, uint32_t s_bot) {
  extern __shared__ uint32_t lds[];
  const uint32_t B = blockDim.x, tid = threadIdx.x;
  const uint32_t i = blockIdx.x * B + tid;
  if (i >= n) return;
  const SceneView& s = tv.top;
  const uint32_t S = s_top + s_bot;
  uint32_t* refs = lds + tid;
  float* ts = reinterpret_cast<float*>(lds + S * B) + tid;
  Stack st{refs, ts, B};
  Stack sb{refs + s_top * B, ts + s_top * B, B};
  uint32_t cnt = 0;
  uint64_t* dst = FILL ? keys + offsets[i] : nullptr;
  const uint32_t root_ref = fbits(s.root[2 * D]);
  if constexpr (MODE == 2) {
    const float* rg = q + (size_t)i * 2 * D;
    if (box_overlap<D>(s.root, s.root + D, rg, rg + D)) {
      walk_visit<D>(s, root_ref,
        [&](const float* nd, bool& ha, bool& hb) {
          if (fbits(nd[4 * D + 3])) { ha = hb = true; return; }   // parts of one reference leaf
          float amn[D], amx[D], bmn[D], bmx[D];
#pragma unroll
          for (int k = 0; k < D; ++k) { amn[k] = nd[4 * k]; amx[k] = nd[4 * k + 1]; bmn[k] = nd[4 * k + 2]; bmx[k] = nd[4 * k + 3]; }
          ha = box_overlap<D>(amn, amx, rg, rg + D);
          hb = box_overlap<D>(bmn, bmx, rg, rg + D);
        },
        [&](uint32_t first, uint32_t c) {
          for (uint32_t j = 0; j < c; ++j) {
            const uint32_t id = s.ids[first + j];
            if (!instance_overlap<D, IK>(tv, id, rg, rg + D)) continue;
            if (FILL) dst[cnt] = id;
            ++cnt;
          }
        }, st);
    }
  } else if constexpr (MODE == 3) {
    CastPolicy<D, KIND_AABB> pol;
    pol.setup(q + (size_t)i * 4 * D);
    if (IK == INST_BALL && pol.root(s.root, max_toi)) {
      walk_visit<D>(s, root_ref,
        [&](const float* nd, bool& ha, bool& hb) {
          if (synthetic_pair<D>(nd)) { ha = hb = true; return; }   // parts of one reference leaf
          float ta, tb;
          pol.pair(nd, max_toi, ha, hb, ta, tb);
        },
        [&](uint32_t first, uint32_t c) {
          for (uint32_t j = 0; j < c; ++j) {
            const uint32_t id = s.ids[first + j];
            float tt;
            if (!instance_cast<D, IK>(tv, pol, id, max_toi, tt)) continue;
            if (FILL) dst[cnt] = hit_key(tt, id);
            ++cnt;
          }
        }, st);
    }
  } else {
    RayD<D> r;
    ray_setup<D>(r, q + (size_t)i * 2 * D);
    float te, tx; int ax;
    if (slab_exact<D>(r, s.root, s.root + D, max_toi, te, tx, ax)) {
      walk_visit<D>(s, root_ref,
        [&](const float* nd, bool& ha, bool& hb) {
          if (synthetic_pair<D>(nd)) { ha = hb = true; return; }   // parts of one reference leaf
          float ta, tb;
          slab_pair<D>(r, nd, max_toi, ha, hb, ta, tb);
        },
        [&](uint32_t first, uint32_t c) {
          for (uint32_t j = 0; j < c; ++j) {
            TlasHit h;
            if (!instance_test<D, IK>(tv, r, s.ids[first + j], max_toi, sb, h)) continue;
            if (FILL) dst[cnt] = hit_key(h.t, h.inst);
            ++cnt;
          }
        }, st);
    }
  }
  if (!FILL) counts[i] = cnt;
}

// Tlas::shapecast_nearest (tlas.rs:306-351) over the probe-grown top-level
// boxes with the shrinking limit; one cast [4D] per thread, the stack in LDS.
// EXACT: unguarded contacts in the reference's order (synthetic pairs passed
// through in leaf order); else guarded contacts, near-first.  Kinds other
// than BALL answer no cast (instance_cast).
template <int D, int IK, bool EXACT>
__global__ __launch_bounds__(1024) void k_tlas_cast_nearest(TlasView tv, const float* __restrict__ q, uint32_t n, float max_toi,
                                                            int32_t* out_inst, float* out_toi, float* out_point,
                                                            uint32_t s_top) {
  extern __shared__ uint32_t lds[];
  const uint32_t B = blockDim.x, tid = threadIdx.x;
  const uint32_t i = blockIdx.x * B + tid;
  if (i >= n) return;
  const SceneView& s = tv.top;
  Stack st{lds + tid, reinterpret_cast<float*>(lds + s_top * B) + tid, B};
  CastPolicy<D, KIND_AABB> pol;
  pol.setup(q + (size_t)i * 4 * D);
  float limit = max_toi, best_t = SQ_INF;
  uint32_t best = NO_LEAF;
  if (IK == INST_BALL && pol.root(s.root, limit)) {
    constexpr int NF = (D + 1) * 4;
    uint32_t cur = fbits(s.root[2 * D]), sp = 0;
    while (true) {
      if (!(cur & LEAF_BIT)) {
        const float* nd = s.nodes + (size_t)cur * NF;
        bool ha, hb, a_first;
        float ta, tb;
        if (EXACT && synthetic_pair<D>(nd)) {
          ha = hb = true; ta = tb = -SQ_INF; a_first = true;
        } else {
          pol.pair(nd, limit, ha, hb, ta, tb);
          a_first = EXACT ? false : ta <= tb;
        }
        const uint32_t ra = fbits(nd[4 * D + 0]), rb = fbits(nd[4 * D + 1]);
        if (ha && hb) {
          st.ref[sp * st.stride] = a_first ? rb : ra;
          st.t[sp * st.stride] = a_first ? tb : ta;
          ++sp;
          cur = a_first ? ra : rb;
          continue;
        }
        if (ha) { cur = ra; continue; }
        if (hb) { cur = rb; continue; }
      } else {
        const uint32_t first = cur & FIRST_MASK;
        const uint32_t cnt = ((cur >> 27) & 15u) + 1u;
        for (uint32_t j = 0; j < cnt; ++j) {
          const uint32_t id = s.ids[first + j];
          float t;
          if (!instance_cast<D, IK, !EXACT>(tv, pol, id, limit, t)) continue;
          if (best == NO_LEAF || hit_before(t, id, best_t, best)) { best = id; best_t = t; limit = t; }
        }
      }
      bool got = false;
      while (sp > 0) {
        --sp;
        if (st.t[sp * st.stride] <= limit) { cur = st.ref[sp * st.stride]; got = true; break; }
      }
      if (!got) break;
    }
  }
  const bool hit = best != NO_LEAF;
  out_inst[i] = hit ? (int32_t)best : -1;
  out_toi[i] = best_t;
  if (out_point) {
#pragma unroll
    for (int k = 0; k < D; ++k) out_point[(size_t)i * D + k] = hit ? pol.r.o[k] + pol.r.d[k] * best_t : 0.0f;
  }
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------


static TlasView tview(const sq_tlas* t) {
  TlasView v;
  v.top = view_of(t->top);
  v.xf = t->xf.as<float>();
  v.world = t->world.as<float>();
  v.blas = t->blas.as<SceneView>();
  v.xs = record_stride(t->in.dim, t->in.kind);
  return v;
}

void tlas_info(const sq_tlas* t, TlasInfo* out) {
  out->ctx = t->ctx;
  out->dim = t->in.dim;
  out->kind = t->in.kind;
  out->n_inst = t->in.n;
  out->s_top = t->top ? t->top->packed.max_stack : 0;
  out->s_bot = t->blas_stack;
  out->view = tview(t);
}

static int tlas_upload_world(sq_tlas* t, std::vector<float>& world) {
  world_boxes(t->in, world);
  return upload(t->world, world.data(), world.size() * 4, t->ctx->stream);
}

static int tlas_upload_records(sq_tlas* t) {
  std::vector<float> rec;
  records(t->in, rec);
  return upload(t->xf, rec.data(), rec.size() * 4, t->ctx->stream);
}

extern "C" {

int sq_tlas_build(sq_ctx* ctx, int dim, int kind, uint64_t n, const float* rot, const float* trans, const float* halves,
                  sq_bvh* const* meshes, uint32_t n_meshes, const uint32_t* mesh_of, const float* scale, sq_tlas** out) {
  if (!ctx || !out || (n && (!rot || !trans))) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = nullptr;
  int rc = dispatch_inst(dim, kind, [&]<int D, int K>() { return SQ_OK; });
  if (rc) return rc;
  if (n > sqh::LEAF_FIRST_MASK) return fail(SQ_ERR_INVALID, "at most 2^27 instances");
  auto t = std::make_unique<sq_tlas>();
  t->ctx = ctx;
  Inst& in = t->in;
  in.dim = dim; in.kind = kind; in.n = n;
  in.rot.assign(rot, rot + n * dim * dim);
  in.trans.assign(trans, trans + n * dim);
  if (kind == INST_BOX) {
    if (n && !halves) return fail(SQ_ERR_INVALID, "box instances need halves");
    in.halves.assign(halves, halves + n * dim);
  } else if (kind == INST_BALL) {
    if (n && !halves) return fail(SQ_ERR_INVALID, "ball instances need radii (in halves)");
    in.halves.assign(halves, halves + n);
    for (uint64_t i = 0; i < n; ++i)
      if (!(in.halves[i] >= 0.0f)) return fail(SQ_ERR_INVALID, "ball radii must be >= 0");
  } else {
    if (n && (!meshes || !mesh_of || !scale)) return fail(SQ_ERR_INVALID, "mesh instances need meshes, mesh_of and scale");
    in.meshes.assign(meshes, meshes + n_meshes);
    for (uint32_t m = 0; m < n_meshes; ++m) {
      if (!meshes[m] || meshes[m]->kind != KIND_TRI || meshes[m]->n_prims == 0 || meshes[m]->ctx != ctx)
        return fail(SQ_ERR_INVALID, "each mesh must be a non-empty TRIANGLE bvh on this context");
      int rm = ensure_host_tree(meshes[m]);             // the mesh root box (local_aabb)
      if (rm) return rm;
    }
    in.mesh_of.assign(mesh_of, mesh_of + n);
    in.scale.assign(scale, scale + n);
    for (uint64_t i = 0; i < n; ++i) {
      if (in.mesh_of[i] >= n_meshes) return fail(SQ_ERR_INVALID, "mesh index out of range");
      if (!(in.scale[i] > 0.0f)) return fail(SQ_ERR_INVALID, "scale must be positive");
    }
    std::vector<SceneView> views(n_meshes);
    for (uint32_t m = 0; m < n_meshes; ++m) {
      views[m] = view_of(meshes[m]);
      t->blas_stack = std::max(t->blas_stack, meshes[m]->packed.max_stack);
    }
    if ((rc = upload(t->blas, views.data(), views.size() * sizeof(SceneView), ctx->stream))) return rc;
  }
  std::vector<float> world;
  if ((rc = tlas_upload_world(t.get(), world))) return rc;
  if ((rc = tlas_upload_records(t.get()))) return rc;
  sq_tree* tree = nullptr;
  if ((rc = sq_tree_build(dim, KIND_AABB, world.data(), n, &tree))) return rc;
  rc = sq_bvh_upload(ctx, tree, KIND_AABB, world.data(), &t->top);
  sq_tree_destroy(tree);
  if (rc) return rc;
  t->baseline = t->top->stats.quality;   // tlas.rs:38: interior surface sum at build
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx_retain(ctx);
  *out = t.release();
  return SQ_OK;
}

int sq_tlas_destroy(sq_tlas* t) {
  if (!t) return SQ_OK;
  sq_ctx* c = t->ctx;
  (void)hipStreamSynchronize(c->stream);
  delete t;
  ctx_release(c);
  return SQ_OK;
}

int sq_tlas_refit(sq_tlas* t, const float* rot, const float* trans, const float* halves) {
  if (!t || (t->in.n && (!rot || !trans))) return fail(SQ_ERR_INVALID, "NULL argument");
  Inst& in = t->in;
  in.rot.assign(rot, rot + in.n * in.dim * in.dim);
  in.trans.assign(trans, trans + in.n * in.dim);
  if (halves && in.kind == INST_BOX) in.halves.assign(halves, halves + in.n * in.dim);
  if (halves && in.kind == INST_BALL) in.halves.assign(halves, halves + in.n);
  if (!in.n) return SQ_OK;
  std::vector<float> world;
  int rc;
  if ((rc = tlas_upload_world(t, world))) return rc;
  if ((rc = tlas_upload_records(t))) return rc;
  // tree::refit over the new world boxes, on the device
  return sq_bvh_refit(t->top, t->world.as<float>());
}

int sq_tlas_download_bounds(sq_tlas* t, float* bounds) {
  if (!t) return fail(SQ_ERR_INVALID, "NULL argument");
  return sq_bvh_download_bounds(t->top, bounds);
}

int sq_tlas_get_stats(const sq_tlas* t, sq_tree_stats* out) {
  if (!t || !out) return fail(SQ_ERR_INVALID, "NULL argument");
  return sq_bvh_get_stats(t->top, out);
}

int sq_tlas_needs_rebuild(sq_tlas* t, float factor, int* out) {
  if (!t || !out) return fail(SQ_ERR_INVALID, "NULL argument");
  *out = 0;
  const size_t nn = t->top->tree.node_count();
  if (!nn) return SQ_OK;
  // tlas.rs:58-60 over the bounds the device currently holds, summed on the
  // host in node order (the sum is order dependent).
  sqh::RefTree cur = t->top->tree;
  int rc = sq_bvh_download_bounds(t->top, cur.bounds.data());
  if (rc) return rc;
  const float q = sqh::tree_interior_surface_sum(cur);
  *out = (t->baseline > 0.0f && q > factor * t->baseline) ? 1 : 0;
  return SQ_OK;
}

static int tlas_nearest_impl(sq_tlas* t, const float* rays, uint64_t n, float max_toi, const float* tmax, NearestOut out, bool any) {
  if (!t || (n && !rays)) return fail(SQ_ERR_INVALID, "NULL argument");
  if (n > 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "at most 2^32-1 rays per call");
  if (!n) return SQ_OK;
  sq_ctx* c = t->ctx;
  if (!t->in.n) {
    return fill_miss(c, n, t->in.dim, out);
  }
  const int B = c->block;
  const uint32_t s_top = t->top->packed.max_stack, s_bot = t->blas_stack;
  TlasView tv = tview(t);
  int rc = begin_timing(c);
  if (rc) return rc;
  if (c->rays_per_lane == 0) {
    // one ray per thread, whole stacks in LDS
    const size_t lds = (size_t)(s_top + s_bot) * B * 8;
    if (lds > 160 * 1024) return fail(SQ_ERR_INVALID, "traversal stacks do not fit in LDS; lower the block size");
    const bool exact = c->hit_mode != SQ_HITS_GUARDED;
    rc = dispatch_inst(t->in.dim, t->in.kind, [&]<int D, int K>() -> int {
      auto go = [&]<bool ANY_, bool EXACT_>() {
        k_tlas_nearest<D, K, ANY_, EXACT_><<<grid_for(n, B), B, lds, c->stream>>>(tv, rays, (uint32_t)n, max_toi, tmax, out, s_top,
                                                                                s_bot);
      };
      if (any) exact ? go.template operator()<true, true>() : go.template operator()<true, false>();
      else exact ? go.template operator()<false, true>() : go.template operator()<false, false>();
      HIP_TRY(hipGetLastError());
      return SQ_OK;
    });
    if (rc) return rc;
    return end_timing(c);
  }
  // pooled, persistent (as nearest_impl): chunks of 64 * per rays, a grid of
  // the resident waves, S-entry rings spilling to a top and a BLAS column
  const uint32_t S = (uint32_t)c->lds_stack;
  const uint32_t per = pooled_per_lane(c, n);
  const uint64_t chunk = 64ull * per;
  const uint64_t chunks = (n + chunk - 1) / chunk;
  const uint32_t rows_top = s_top > S ? s_top - S : 0, rows_bot = s_bot > S ? s_bot - S : 0;
  const size_t lds = (size_t)2 * S * B * 8;
  rc = dispatch_inst(t->in.dim, t->in.kind, [&]<int D, int K>() -> int {
    auto launch = [&]<bool ANY_, bool EXACT_>() -> int {
      int per_cu = 0;
      if (int r0 = blocks_per_cu(reinterpret_cast<const void*>(&k_tlas_nearest_pool<D, K, ANY_, EXACT_>), c->device, B, lds, &per_cu))
        return r0;
      const uint64_t wpb = (uint64_t)B / 64;
      const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)c->cus * std::max(1, per_cu), (chunks + wpb - 1) / wpb));
      const uint64_t lanes = grid * (uint64_t)B;
      int r2 = c->spill.ensure((size_t)(rows_top + rows_bot) * lanes * sizeof(SpillEntry) + 64);
      if (r2) return r2;
      SpillEntry* sp = c->spill.as<SpillEntry>();
      k_tlas_nearest_pool<D, K, ANY_, EXACT_><<<(uint32_t)grid, B, lds, c->stream>>>(
          tv, rays, (uint32_t)n, max_toi, tmax, out, (uint32_t)chunk, c->pool_ctr.as<uint32_t>(), S, sp,
          sp + (size_t)rows_top * lanes, (uint32_t)lanes, (uint32_t)c->descend_min);
      HIP_TRY(hipGetLastError());
      return SQ_OK;
    };
    if (c->hit_mode == SQ_HITS_GUARDED) return any ? launch.template operator()<true, false>() : launch.template operator()<false, false>();
    return any ? launch.template operator()<true, true>() : launch.template operator()<false, true>();
  });
  if (rc) return rc;
  return end_timing(c);
}

int sq_tlas_raycast_nearest(sq_tlas* t, const float* rays, uint64_t n, float max_toi, const float* tmax, int32_t* out_inst,
                            float* out_toi, float* out_point, float* out_normal, int32_t* out_sub) {
  if (n && (!out_inst || !out_toi)) return fail(SQ_ERR_INVALID, "out_inst and out_toi are required");
  NearestOut o{out_inst, out_toi, out_point, out_normal, out_sub, nullptr};
  return tlas_nearest_impl(t, rays, n, max_toi, tmax, o, false);
}

int sq_tlas_raycast_any(sq_tlas* t, const float* rays, uint64_t n, float max_toi, const float* tmax, uint8_t* out_hit) {
  if (n && !out_hit) return fail(SQ_ERR_INVALID, "out_hit is required");
  NearestOut o{nullptr, nullptr, nullptr, nullptr, nullptr, out_hit};
  return tlas_nearest_impl(t, rays, n, max_toi, tmax, o, true);
}

static int tlas_collect_impl(sq_tlas* t, const float* q, uint64_t n, float max_toi, int mode, uint64_t* offsets,
                             int32_t* out_inst, float* out_toi, uint64_t capacity, uint64_t* total) {
  if (!t || !offsets || !total || (n && !q)) return fail(SQ_ERR_INVALID, "NULL argument");
  if (n > 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "at most 2^32-1 queries per call");
  sq_ctx* c = t->ctx;
  *total = 0;
  int rc;
  if (!n || !t->in.n) {
    HIP_TRY(hipMemsetAsync(offsets, 0, (n + 1) * 8, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SQ_OK;
  }
  if ((rc = c->counts.ensure(n * 4))) return rc;
  const int B = c->block;
  const uint32_t s_top = t->top->packed.max_stack, s_bot = t->blas_stack;
  const size_t lds = (size_t)(s_top + s_bot) * B * 8;
  if (lds > 160 * 1024) return fail(SQ_ERR_INVALID, "traversal stacks do not fit in LDS; lower the block size");
  TlasView tv = tview(t);
  uint32_t* counts = c->counts.as<uint32_t>();
  auto launch = [&](bool fill, uint64_t* keys) -> int {
    return dispatch_inst(t->in.dim, t->in.kind, [&]<int D, int K>() -> int {
      dim3 g(grid_for(n, B)), b(B);
      const uint32_t nn = (uint32_t)n;
      if (mode == 2) {
        if (fill) k_tlas_collect<D, K, 2, true><<<g, b, lds, c->stream>>>(tv, q, nn, max_toi, counts, offsets, keys, s_top, s_bot);
        else k_tlas_collect<D, K, 2, false><<<g, b, lds, c->stream>>>(tv, q, nn, max_toi, counts, offsets, keys, s_top, s_bot);
      } else if (mode == 3) {
        if (fill) k_tlas_collect<D, K, 3, true><<<g, b, lds, c->stream>>>(tv, q, nn, max_toi, counts, offsets, keys, s_top, s_bot);
        else k_tlas_collect<D, K, 3, false><<<g, b, lds, c->stream>>>(tv, q, nn, max_toi, counts, offsets, keys, s_top, s_bot);
      } else {
        if (fill) k_tlas_collect<D, K, 0, true><<<g, b, lds, c->stream>>>(tv, q, nn, max_toi, counts, offsets, keys, s_top, s_bot);
        else k_tlas_collect<D, K, 0, false><<<g, b, lds, c->stream>>>(tv, q, nn, max_toi, counts, offsets, keys, s_top, s_bot);
      }
      HIP_TRY(hipGetLastError());
      return SQ_OK;
    });
  };
  return csr_run(c, n, offsets, out_inst, out_toi, mode != 2, capacity, total, launch);
}

int sq_tlas_raycast_all(sq_tlas* t, const float* rays, uint64_t n, float max_toi, uint64_t* offsets, int32_t* out_inst,
                        float* out_toi, uint64_t capacity, uint64_t* total) {
  return tlas_collect_impl(t, rays, n, max_toi, 0, offsets, out_inst, out_toi, capacity, total);
}

int sq_tlas_overlap(sq_tlas* t, const float* regions, uint64_t n, uint64_t* offsets, int32_t* out_inst, uint64_t capacity,
                    uint64_t* total) {
  return tlas_collect_impl(t, regions, n, 0.0f, 2, offsets, out_inst, nullptr, capacity, total);
}

int sq_tlas_shapecast_all(sq_tlas* t, const float* casts, uint64_t n, float max_toi, uint64_t* offsets, int32_t* out_inst,
                          float* out_toi, uint64_t capacity, uint64_t* total) {
  return tlas_collect_impl(t, casts, n, max_toi, 3, offsets, out_inst, out_toi, capacity, total);
}

int sq_tlas_shapecast_nearest(sq_tlas* t, const float* casts, uint64_t n, float max_toi, int32_t* out_inst, float* out_toi,
                              float* out_point) {
  if (!t || (n && (!casts || !out_inst || !out_toi))) return fail(SQ_ERR_INVALID, "NULL argument");
  if (n > 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "at most 2^32-1 casts per call");
  if (!n) return SQ_OK;
  sq_ctx* c = t->ctx;
  if (!t->in.n) return fill_miss(c, n, t->in.dim, NearestOut{out_inst, out_toi, out_point, nullptr, nullptr, nullptr});
  const int B = c->block;
  const uint32_t s_top = t->top->packed.max_stack;
  const size_t lds = (size_t)s_top * B * 8;
  if (lds > 160 * 1024) return fail(SQ_ERR_INVALID, "traversal stack does not fit in LDS; lower the block size");
  TlasView tv = tview(t);
  int rc = begin_timing(c);
  if (rc) return rc;
  const bool exact = c->hit_mode != SQ_HITS_GUARDED;
  rc = dispatch_inst(t->in.dim, t->in.kind, [&]<int D, int K>() -> int {
    if (exact)
      k_tlas_cast_nearest<D, K, true><<<grid_for(n, B), B, lds, c->stream>>>(tv, casts, (uint32_t)n, max_toi, out_inst, out_toi,
                                                                           out_point, s_top);
    else
      k_tlas_cast_nearest<D, K, false><<<grid_for(n, B), B, lds, c->stream>>>(tv, casts, (uint32_t)n, max_toi, out_inst, out_toi,
                                                                            out_point, s_top);
    HIP_TRY(hipGetLastError());
    return SQ_OK;
  });
  if (rc) return rc;
  return end_timing(c);
}

}  // extern "C"

sq_bvh* tlas_top_tree(sq_tlas* t) { return t->top; }
int tlas_dim(const sq_tlas* t) { return t->in.dim; }

int tlas_nearest_proposed(sq_tlas* t, const float* rays, uint64_t n, float max_toi, const uint16_t* top, const uint32_t* sub_ref,
                          uint32_t n_sub, int32_t* out_inst, float* out_toi, int32_t* out_sub, uint64_t* stats) {
  if (n > 0xFFFFFFFFull) return fail(SQ_ERR_INVALID, "at most 2^32-1 rays per call");
  if (t->in.dim != 3) return fail(SQ_ERR_UNSUPPORTED, "the proposal serves D = 3 scenes");
  if (!n) return SQ_OK;
  sq_ctx* c = t->ctx;
  if (!t->in.n) return fill_miss(c, n, 3, NearestOut{out_inst, out_toi, nullptr, nullptr, out_sub, nullptr});
  const int B = std::min(c->block, 256);
  const uint32_t s_top = t->top->packed.max_stack, s_bot = t->blas_stack;
  const uint32_t S = (uint32_t)c->lds_stack;
  const uint32_t per = pooled_per_lane(c, n);
  const uint64_t chunk = 64ull * per;
  const uint64_t chunks = (n + chunk - 1) / chunk;
  const uint32_t rows_top = s_top > S ? s_top - S : 0, rows_bot = s_bot > S ? s_bot - S : 0;
  const size_t lds = (size_t)2 * S * B * 8;
  TlasView tv = tview(t);
  int rc = begin_timing(c);
  if (rc) return rc;
  auto launch = [&]<int K, bool COUNT>() -> int {
    int per_cu = 0;
    if (int r0 = blocks_per_cu(reinterpret_cast<const void*>(&k_tlas_nearest_proposed<K, COUNT>), c->device, B, lds, &per_cu))
      return r0;
    const uint64_t wpb = (uint64_t)B / 64;
    const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)c->cus * std::max(1, per_cu), (chunks + wpb - 1) / wpb));
    const uint64_t lanes = grid * (uint64_t)B;
    int r2 = c->spill.ensure((size_t)(rows_top + rows_bot) * lanes * sizeof(SpillEntry) + 64);
    if (r2) return r2;
    SpillEntry* sp = c->spill.as<SpillEntry>();
    k_tlas_nearest_proposed<K, COUNT><<<(uint32_t)grid, B, lds, c->stream>>>(
        tv, rays, (uint32_t)n, max_toi, top, sub_ref, n_sub, out_inst, out_toi, out_sub, (uint32_t)chunk, c->pool_ctr.as<uint32_t>(),
        S, sp, sp + (size_t)rows_top * lanes, (uint32_t)lanes, (uint32_t)c->descend_min,
        reinterpret_cast<unsigned long long*>(stats));
    HIP_TRY(hipGetLastError());
    return SQ_OK;
  };
  auto by_count = [&]<int K>() -> int {
    return stats ? launch.template operator()<K, true>() : launch.template operator()<K, false>();
  };
  switch (t->in.kind) {
    case INST_BOX: rc = by_count.template operator()<INST_BOX>(); break;
    case INST_MESH: rc = by_count.template operator()<INST_MESH>(); break;
    case INST_BALL: rc = by_count.template operator()<INST_BALL>(); break;
    default: rc = fail(SQ_ERR_UNSUPPORTED, "instance kind");
  }
  if (rc) return rc;
  return end_timing(c);
}
