// emu.cpp -- TEST-ONLY CPU build of the engine's per-query logic.
//
// Compiles spatial-query_amd/csrc/traverse.h and host_tree.cpp with g++ and
// runs the exact walks the HIP kernels run (same packed tree, same leaf-ordered
// records), one query per loop iteration with a private stack.  It lets the
// test suite check the kernel logic against the oracle on the CPU; it is
// never part of the product and never a fallback for it.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../spatial-query_amd/csrc/host_tree.h"
#include "../../spatial-query_amd/csrc/traverse.h"

using namespace sqd;

namespace {

struct Built {
  sqh::RefTree tree;
  sqh::Packed packed;
  std::vector<float> prims;
  std::vector<float> quads;     // 4-wide records (quad_records), built on first use
  SceneView view() const {
    SceneView s;
    s.nodes = packed.nodes.data();
    s.prims = prims.data();
    s.ids = tree.prim_indices.data();
    s.root = packed.root.data();
    s.quads = quads.empty() ? nullptr : quads.data();
    return s;
  }
};

template <int D> void build_quads(Built& b) {
  constexpr int NF = (D + 1) * 4;
  const uint32_t n = (uint32_t)(b.packed.nodes.size() / NF);
  b.quads.assign((size_t)std::max<uint32_t>(n, 1) * 2 * NF, 0.0f);
  for (uint32_t p = 0; p < n; ++p) quad_records<D>(b.packed.nodes.data(), n, p, b.quads.data() + (size_t)p * 2 * NF);
}

void build(int dim, int kind, const float* prims, uint32_t n, Built& b) {
  std::vector<float> boxes;
  sqh::prim_boxes(dim, kind, prims, n, boxes);
  sqh::build_tree(dim, boxes.data(), n, b.tree);
  sqh::pack_tree(b.tree, b.packed);
  sqh::pack_prims(dim, kind, prims, b.tree, b.prims);
}

template <class F> void par_for(uint64_t n, int threads, F&& f) {
  if (threads <= 1 || n < 1024) { for (uint64_t i = 0; i < n; ++i) f(i); return; }
  std::vector<std::thread> th;
  uint64_t chunk = (n + threads - 1) / threads;
  for (uint64_t lo = 0; lo < n; lo += chunk)
    th.emplace_back([&, lo] { for (uint64_t i = lo; i < std::min(n, lo + chunk); ++i) f(i); });
  for (auto& t : th) t.join();
}

// The pooled kernel's per-lane walk: walk_begin + walk_step with a ring stack
// of `ring` LDS slots (here a local array) spilling to a bounds-checked column.
template <int D, int K, bool QUAD = false, bool REF = false, class Pol>
NearestResult ring_walk(const Built& b, const Pol& pol, float max_toi, int any, uint32_t ring, uint32_t* prims_tested = nullptr) {
  constexpr int PS = Prim<D, K>::STRIDE;
  const SceneView s = b.view();
  Filter f{nullptr, ~0ull, -1};
  std::vector<uint32_t> lref(ring);
  std::vector<float> lt(ring);
  // the engine's stack bound: pair walk max_stack; 4-wide walk 3 (max_stack / 2 + 1) + 1
  const uint32_t bound = QUAD ? 3 * (b.packed.max_stack / 2 + 1) + 1 : b.packed.max_stack;
  const uint32_t rows = bound > ring ? bound - ring : 0;
  std::vector<SpillEntry> col(rows + 1);
  RingStack st{lref.data(), lt.data(), 1, ring, col.data(), 1};
  WalkState w;
  if (!walk_begin<D>(s, pol, max_toi, w)) return w.res;
  WorkCount wc;
  for (;;) {
    const bool done = any ? walk_step<D, PS, true, false, true, false, QUAD>(s, pol, f, st, w, &wc)
                          : walk_step<D, PS, false, false, true, false, QUAD, REF>(s, pol, f, st, w, &wc);
    if (w.sp > bound) { fprintf(stderr, "stack bound exceeded\n"); abort(); }
    if (done) {
      if (prims_tested) *prims_tested = wc.prims;
      return w.res;
    }
  }
}

// GUARD: the engine's SQ_HITS_GUARDED walk (guarded kinds, near-first); else
// SQ_HITS_REFERENCE (unguarded, nearest queries over leaves whose hits may
// leave their box in the reference's order).
template <int D, int K, bool GUARD>
void nearest_dk(const Built& b, const float* q, uint64_t nq, float max_toi, int any, int cast, int threads,
                int32_t* leaf, float* toi, uint32_t ring) {
  constexpr int PS = Prim<D, K>::STRIDE;
  using RP = RayPolicy<D, K, GUARD>;
  using CP = CastPolicy<D, K, float, GUARD>;
  constexpr bool REF_R = !GUARD && !RP::INSIDE, REF_C = !GUARD && !CP::INSIDE;
  const SceneView s = b.view();
  const uint32_t S = b.packed.max_stack;
  Filter f{nullptr, ~0ull, -1};
  if (ring) {
    par_for(nq, threads, [&](uint64_t i) {
      NearestResult h;
      if (cast) { CP pol; pol.setup(q + i * 4 * D); h = ring_walk<D, K, false, REF_C>(b, pol, max_toi, 0, ring); }
      else { RP pol; pol.setup(q + i * 2 * D); h = ring_walk<D, K, false, REF_R>(b, pol, max_toi, any, ring); }
      leaf[i] = h.leaf == NO_LEAF ? -1 : (int32_t)h.leaf;
      toi[i] = h.t;
    });
    return;
  }
  par_for(nq, threads, [&](uint64_t i) {
    std::vector<uint32_t> sr(S);
    std::vector<float> stt(S);
    Stack st{sr.data(), stt.data(), 1};
    NearestResult h;
    if (cast) {
      CP pol; pol.setup(q + i * 4 * D);
      h = walk_nearest<D, PS, false, false, REF_C>(s, pol, max_toi, f, st);
    } else {
      RP pol; pol.setup(q + i * 2 * D);
      if (any) h = walk_nearest<D, PS, true, false>(s, pol, max_toi, f, st);
      else h = walk_nearest<D, PS, false, false, REF_R>(s, pol, max_toi, f, st);
    }
    leaf[i] = h.leaf == NO_LEAF ? -1 : (int32_t)h.leaf;
    toi[i] = h.t;
  });
}

// The 4-wide walk (walk_step with QUAD) of every ray, ring stack as above:
// GUARD near-first over the guarded kinds; else unguarded, nearest queries
// over leaves whose hits may leave their box in the reference's order (the
// engine takes this form only for trees without synthetic pairs).
template <int D, int K, bool GUARD>
void nearest_quad_dk(Built& b, const float* q, uint64_t nq, float max_toi, int any, int threads, int32_t* leaf, float* toi,
                     uint32_t ring, uint32_t* prims_tested) {
  if (b.quads.empty()) build_quads<D>(b);
  constexpr bool REF = !GUARD && !RayPolicy<D, K, GUARD>::INSIDE;
  par_for(nq, threads, [&](uint64_t i) {
    RayPolicy<D, K, GUARD> pol; pol.setup(q + i * 2 * D);
    NearestResult h = ring_walk<D, K, true, REF>(b, pol, max_toi, any, ring, prims_tested ? prims_tested + i : nullptr);
    leaf[i] = h.leaf == NO_LEAF ? -1 : (int32_t)h.leaf;
    toi[i] = h.t;
  });
}

template <int D>
void required_d(const Built& b, const float* rays, uint64_t nq, const float* tl, uint32_t* on, uint32_t* op) {
  const SceneView s = b.view();
  const uint32_t S = b.packed.max_stack;
  for (uint64_t i = 0; i < nq; ++i) {
    std::vector<uint32_t> sr(S);
    Stack st{sr.data(), nullptr, 1};
    RayD<D> r; ray_setup<D>(r, rays + i * 2 * D);
    walk_required<D>(s, r, tl[i], st, on[i], op[i]);
  }
}

template <class F> int dk(int dim, int kind, F&& f) {
  if (kind == 2) { f.template operator()<3, 2>(); return 0; }
  switch (dim) {
    case 2: kind ? f.template operator()<2, 1>() : f.template operator()<2, 0>(); return 0;
    case 3: kind ? f.template operator()<3, 1>() : f.template operator()<3, 0>(); return 0;
    case 4: kind ? f.template operator()<4, 1>() : f.template operator()<4, 0>(); return 0;
    case 5: kind ? f.template operator()<5, 1>() : f.template operator()<5, 0>(); return 0;
    case 6: kind ? f.template operator()<6, 1>() : f.template operator()<6, 0>(); return 0;
  }
  return 1;
}

}  // namespace

extern "C" {

void* emu_build(int dim, int kind, const float* prims, uint32_t n) {
  Built* b = new Built();
  build(dim, kind, prims, n, *b);
  return b;
}
void emu_free(void* h) { delete (Built*)h; }
uint32_t emu_max_stack(void* h) { return ((Built*)h)->packed.max_stack; }
uint32_t emu_n_pairs(void* h) { return ((Built*)h)->packed.n_pairs; }

int emu_nearest(void* h, int dim, int kind, const float* q, uint64_t nq, float max_toi, int any, int cast, int threads,
                int32_t* leaf, float* toi, uint32_t ring, int guarded) {
  const Built& b = *(Built*)h;
  if (b.tree.prim_indices.empty()) { for (uint64_t i = 0; i < nq; ++i) { leaf[i] = -1; toi[i] = SQ_INF; } return 0; }
  return dk(dim, kind, [&]<int D, int K>() {
    if (guarded) nearest_dk<D, K, true>(b, q, nq, max_toi, any, cast, threads, leaf, toi, ring);
    else nearest_dk<D, K, false>(b, q, nq, max_toi, any, cast, threads, leaf, toi, ring);
  });
}

// The 4-wide walk (tuning `wide`) with a ring stack of `ring` slots; -2 when
// an unguarded walk is asked of a tree with synthetic pairs.
// Pauses every `every`-th node (resumable descents, as a wave pauses them).
struct EveryPause {
  uint32_t every;
  mutable uint32_t n = 0;
  bool operator()() const { return every && ++n % every == 0; }
};

// The look-ahead walk in the reference's order (walk_step_spec) of every ray
// over the 4-wide records (quad) or the pair nodes, unguarded, with a ring
// stack and a descent paused every `pause_every` nodes (0: never); returns
// the primitives each walk tested.  -2 for quad over a tree with synthetic
// pairs.
int emu_nearest_spec(void* h, int dim, int kind, const float* q, uint64_t nq, float max_toi, int quad, uint32_t ring,
                     uint32_t pause_every, int32_t* leaf, float* toi, uint32_t* prims_tested) {
  Built& b = *(Built*)h;
  if (b.tree.prim_indices.empty() || !ring || kind == 0) return -1;
  if (quad && b.packed.n_synth != 0) return -2;
  return dk(dim, kind, [&]<int D, int K>() {
    if (quad && b.quads.empty()) build_quads<D>(b);
    constexpr int PS = Prim<D, K>::STRIDE;
    const SceneView s = b.view();
    const uint32_t bound = quad ? 3 * (b.packed.max_stack / 2 + 1) + 1 : b.packed.max_stack;
    par_for(nq, 8, [&](uint64_t i) {
      RayPolicy<D, K, false> pol; pol.setup(q + i * 2 * D);
      Filter f{nullptr, ~0ull, -1};
      std::vector<uint32_t> lref(ring);
      std::vector<float> lt(ring);
      const uint32_t rows = bound > ring ? bound - ring : 0;
      std::vector<SpillEntry> col(rows + 1);
      RingStack st{lref.data(), lt.data(), 1, ring, col.data(), 1};
      SpecState w;
      WorkCount wc;
      if (spec_begin<D>(s, pol, max_toi, w)) {
        EveryPause pz{pause_every};
        for (;;) {
          const bool done = quad ? walk_step_spec<D, PS, false, true, RayPolicy<D, K, false>, EveryPause, true>(s, pol, f, st, w, pz, &wc)
                                 : walk_step_spec<D, PS, false, false, RayPolicy<D, K, false>, EveryPause, true>(s, pol, f, st, w, pz, &wc);
          if (w.sp > bound) { fprintf(stderr, "stack bound exceeded\n"); abort(); }
          if (done) break;
        }
      }
      leaf[i] = w.res.leaf == NO_LEAF ? -1 : (int32_t)w.res.leaf;
      toi[i] = w.res.t;
      prims_tested[i] = wc.prims;
    });
  });
}

// prims_tested (optional, [nq]): primitives each walk tested.
int emu_nearest_quad(void* h, int dim, int kind, const float* q, uint64_t nq, float max_toi, int any, int threads,
                     int32_t* leaf, float* toi, uint32_t ring, int guarded, uint32_t* prims_tested) {
  Built& b = *(Built*)h;
  if (b.tree.prim_indices.empty()) { for (uint64_t i = 0; i < nq; ++i) { leaf[i] = -1; toi[i] = SQ_INF; } return 0; }
  if (!guarded && kind != 0 && b.packed.n_synth != 0) return -2;
  return dk(dim, kind, [&]<int D, int K>() {
    if (guarded) nearest_quad_dk<D, K, true>(b, q, nq, max_toi, any, threads, leaf, toi, ring, prims_tested);
    else nearest_quad_dk<D, K, false>(b, q, nq, max_toi, any, threads, leaf, toi, ring, prims_tested);
  });
}

// The pair walk of emu_nearest with a ring stack (ring > 0), returning the
// primitives each walk tested as well.
int emu_nearest_pair_counted(void* h, int dim, int kind, const float* q, uint64_t nq, float max_toi, int guarded, uint32_t ring,
                             int32_t* leaf, float* toi, uint32_t* prims_tested) {
  const Built& b = *(Built*)h;
  if (b.tree.prim_indices.empty() || !ring) return -1;
  return dk(dim, kind, [&]<int D, int K>() {
    auto go = [&]<bool GUARD>() {
      constexpr bool REF = !GUARD && !RayPolicy<D, K, GUARD>::INSIDE;
      par_for(nq, 8, [&](uint64_t i) {
        RayPolicy<D, K, GUARD> pol; pol.setup(q + i * 2 * D);
        NearestResult r = ring_walk<D, K, false, REF>(b, pol, max_toi, 0, ring, prims_tested + i);
        leaf[i] = r.leaf == NO_LEAF ? -1 : (int32_t)r.leaf;
        toi[i] = r.t;
      });
    };
    if (guarded) go.template operator()<true>();
    else go.template operator()<false>();
  });
}
uint32_t emu_n_synth(void* h) { return ((Built*)h)->packed.n_synth; }

// The reference's counters of every ray (walk_counted, as k_counted runs it):
// nearest (all = 0, leaf and toi written) or all hits; stats[nq][4].
int emu_counted(void* h, int dim, int kind, const float* q, uint64_t nq, float max_toi, int all, int32_t* leaf, float* toi,
                uint32_t* stats) {
  const Built& b = *(Built*)h;
  if (b.tree.prim_indices.empty()) {
    for (uint64_t i = 0; i < nq; ++i) { leaf[i] = -1; toi[i] = SQ_INF; for (int k = 0; k < 4; ++k) stats[4 * i + k] = 0; }
    return 0;
  }
  return dk(dim, kind, [&]<int D, int K>() {
    const SceneView s = b.view();
    const uint32_t S = b.packed.max_stack;
    const Filter f{nullptr, ~0ull, -1};
    par_for(nq, 8, [&](uint64_t i) {
      std::vector<uint32_t> sr(S);
      std::vector<float> stt(S);
      RayPolicy<D, K, false> pol;
      pol.setup(q + i * 2 * D);
      NearestResult hres = all ? walk_counted<D, Prim<D, K>::STRIDE, true, false>(s, pol, max_toi, f, Stack{sr.data(), stt.data(), 1}, stats + 4 * i)
                               : walk_counted<D, Prim<D, K>::STRIDE, false, false>(s, pol, max_toi, f, Stack{sr.data(), stt.data(), 1}, stats + 4 * i);
      leaf[i] = hres.leaf == NO_LEAF ? -1 : (int32_t)hres.leaf;
      toi[i] = hres.t;
    });
  });
}

// mode 0 all, 1 crossings, 2 overlap, 3 shapecast-all.  Returns the total;
// writes counts[nq] and, when keys != NULL, the unsorted keys per query at
// their CSR offsets (offsets computed here).
uint64_t emu_collect(void* h, int dim, int kind, const float* q, uint64_t nq, float max_toi, int mode,
                     uint64_t* offsets, uint64_t* keys) {
  const Built& b = *(Built*)h;
  uint64_t total = 0;
  offsets[0] = 0;
  if (b.tree.prim_indices.empty()) { for (uint64_t i = 0; i < nq; ++i) offsets[i + 1] = 0; return 0; }
  const SceneView s = b.view();
  const uint32_t S = b.packed.max_stack;
  Filter f{nullptr, ~0ull, -1};
  dk(dim, kind, [&]<int D, int K>() {
    constexpr int PS = Prim<D, K>::STRIDE;
    for (uint64_t i = 0; i < nq; ++i) {
      std::vector<uint32_t> sr(S);
      Stack st{sr.data(), nullptr, 1};
      std::vector<uint64_t> out;
      const uint32_t root_ref = fbits(s.root[2 * D]);
      if (mode == 2) {
        const float* rg = q + i * 2 * D;
        if (box_overlap<D>(s.root, s.root + D, rg, rg + D))
          walk_visit<D>(s, root_ref,
            [&](const float* nd, bool& ha, bool& hb) {
              if (K != KIND_AABB && synthetic_pair<D>(nd)) { ha = hb = true; return; }
              float amn[D], amx[D], bmn[D], bmx[D];
              for (int k = 0; k < D; ++k) { amn[k] = nd[4 * k]; amx[k] = nd[4 * k + 1]; bmn[k] = nd[4 * k + 2]; bmx[k] = nd[4 * k + 3]; }
              ha = box_overlap<D>(amn, amx, rg, rg + D); hb = box_overlap<D>(bmn, bmx, rg, rg + D);
            },
            [&](uint32_t first, uint32_t c) {
              for (uint32_t j = 0; j < c; ++j)
                if (overlap_narrow<D, K>(s.prims + (size_t)(first + j) * PS, rg, rg + D)) out.push_back(s.ids[first + j]);
            }, st);
      } else {
        auto run = [&](auto& pol) {
          if (!pol.root(s.root, max_toi)) return;
          walk_visit<D>(s, root_ref,
            [&](const float* nd, bool& ha, bool& hb) {
              if (K != KIND_AABB && synthetic_pair<D>(nd)) { ha = hb = true; return; }
              float ta, tb;
              pol.pair(nd, max_toi, ha, hb, ta, tb);
            },
            [&](uint32_t first, uint32_t c) {
              for (uint32_t j = 0; j < c; ++j) {
                const float* p = s.prims + (size_t)(first + j) * PS;
                float t;
                if constexpr (K == KIND_AABB) {
                  if (mode == 1) {
                    float te, tx; int ax;
                    if (!slab_exact<D>(pol.r, p, p + D, max_toi, te, tx, ax)) continue;
                    out.push_back(hit_key(te, s.ids[first + j]));
                    if (tx > te) {
                      float te2, tx2;
                      slab_exact<D>(pol.r, p, p + D, SQ_INF, te2, tx2, ax);
                      if (tx2 <= max_toi) out.push_back(hit_key(tx2, s.ids[first + j]));
                    }
                    continue;
                  }
                }
                if (pol.narrow(p, max_toi, t)) out.push_back(hit_key(t, s.ids[first + j]));
              }
            }, st);
        };
        if (mode == 3) { CastPolicy<D, K> pol; pol.setup(q + i * 4 * D); run(pol); }
        else { RayPolicy<D, K> pol; pol.setup(q + i * 2 * D); run(pol); }
      }
      offsets[i + 1] = offsets[i] + out.size();
      if (keys) std::memcpy(keys + offsets[i], out.data(), out.size() * 8);
    }
    total = offsets[nq];
  });
  return total;
}

int emu_required(void* h, int dim, const float* rays, uint64_t nq, const float* tl, uint32_t* on, uint32_t* op) {
  const Built& b = *(Built*)h;
  switch (dim) {
    case 2: required_d<2>(b, rays, nq, tl, on, op); return 0;
    case 3: required_d<3>(b, rays, nq, tl, on, op); return 0;
    case 4: required_d<4>(b, rays, nq, tl, on, op); return 0;
    case 5: required_d<5>(b, rays, nq, tl, on, op); return 0;
    case 6: required_d<6>(b, rays, nq, tl, on, op); return 0;
  }
  return 1;
}

float emu_key_toi(uint64_t k) { return key_toi(k); }

}  // extern "C"
