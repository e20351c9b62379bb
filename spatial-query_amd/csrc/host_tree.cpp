// host_tree.cpp -- the reference median-split build, parallel over subtrees,
// and the re-pack into the device pair-node layout.
//
// The build reproduces crates/spatial-query/src/accel/tree.rs:112-163 exactly
// (same node order, same bounds, same prim_indices), because the device walks
// the reference's tree.  It differs only in how it gets there: centroid keys
// are precomputed as total-order integers, each node sorts (key, index) pairs
// (a stable sort by key == the reference's stable sort_by(total_cmp)), and the
// two halves of large nodes are built on separate threads into local node
// arrays that are spliced back in pre-order.
#include "host_tree.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <future>
#include <stdexcept>

namespace sqh {

namespace {

constexpr uint32_t LEAF_SIZE = 4;                  // tree.rs:12
template <class R> struct Eps;
template <> struct Eps<float> { static constexpr float v = 1.1920929e-7f; };            // f32::EPSILON
template <> struct Eps<double> { static constexpr double v = 2.220446049250313e-16; };  // f64::EPSILON
constexpr uint32_t PAR_MIN_COUNT = 1u << 15;       // split work across threads above this
constexpr int PAR_MAX_DEPTH = 4;                   // at most 2^4 concurrent subtrees

// f32::total_cmp / f64::total_cmp order as an unsigned key.
inline uint32_t total_ukey(float f) {
  int32_t i;
  std::memcpy(&i, &f, 4);
  i ^= (int32_t)(((uint32_t)(i >> 31)) >> 1);
  return (uint32_t)i ^ 0x80000000u;
}
inline uint64_t total_ukey(double f) {
  int64_t i;
  std::memcpy(&i, &f, 8);
  i ^= (int64_t)(((uint64_t)(i >> 63)) >> 1);
  return (uint64_t)i ^ 0x8000000000000000ull;
}

template <class R> struct Ctx {
  using Key = decltype(total_ukey(R(0)));
  int dim;
  const R* boxes;                     // [n][2D]
  std::vector<R> cent;                // [n][D]
  std::vector<Key> keys;              // [D][n]
  uint32_t n;
};

template <class R> struct Local {
  std::vector<R> bounds;
  std::vector<uint32_t> ca, cb, pc;
  uint32_t push(int dim, const R* b, uint32_t a, uint32_t c, uint32_t p) {
    bounds.insert(bounds.end(), b, b + 2 * dim);
    ca.push_back(a); cb.push_back(c); pc.push_back(p);
    return (uint32_t)ca.size() - 1;
  }
};

// Stable order of idx[start, end) by key: (key, original position) order.
inline void sort_by_key(const uint32_t* key, uint32_t* idx, uint32_t start, uint32_t count) {
  std::vector<uint64_t> kv(count);
  for (uint32_t i = 0; i < count; ++i) kv[i] = ((uint64_t)key[idx[start + i]] << 32) | i;
  std::sort(kv.begin(), kv.end());
  std::vector<uint32_t> tmp(count);
  for (uint32_t i = 0; i < count; ++i) tmp[i] = idx[start + (uint32_t)kv[i]];
  std::memcpy(idx + start, tmp.data(), count * sizeof(uint32_t));
}
inline void sort_by_key(const uint64_t* key, uint32_t* idx, uint32_t start, uint32_t count) {
  std::vector<std::pair<uint64_t, uint32_t>> kv(count);
  for (uint32_t i = 0; i < count; ++i) kv[i] = {key[idx[start + i]], i};
  std::sort(kv.begin(), kv.end());
  std::vector<uint32_t> tmp(count);
  for (uint32_t i = 0; i < count; ++i) tmp[i] = idx[start + kv[i].second];
  std::memcpy(idx + start, tmp.data(), count * sizeof(uint32_t));
}

template <class R>
uint32_t build_range(const Ctx<R>& cx, uint32_t* idx, uint32_t start, uint32_t end, int depth, Local<R>& out) {
  const int D = cx.dim;
  // tree.rs:120-123: fold union over the range, from EMPTY.
  R nb[2 * MAX_DIM];
  for (int k = 0; k < D; ++k) { nb[k] = INFINITY; nb[D + k] = -INFINITY; }
  for (uint32_t i = start; i < end; ++i) {
    const R* b = cx.boxes + (size_t)idx[i] * 2 * D;
    for (int k = 0; k < D; ++k) { nb[k] = std::fmin(nb[k], b[k]); nb[D + k] = std::fmax(nb[D + k], b[D + k]); }
  }
  const uint32_t count = end - start;
  const uint32_t node = out.push(D, nb, start, 0, count);   // :129-134
  if (count <= LEAF_SIZE) return node;                       // :136-138
  // :140-148: centroid bounds, longest axis (first max wins), coincidence stop.
  R cmin[MAX_DIM], cmax[MAX_DIM];
  for (int k = 0; k < D; ++k) { cmin[k] = INFINITY; cmax[k] = -INFINITY; }
  for (uint32_t i = start; i < end; ++i) {
    const R* c = cx.cent.data() + (size_t)idx[i] * D;
    for (int k = 0; k < D; ++k) { cmin[k] = std::fmin(cmin[k], c[k]); cmax[k] = std::fmax(cmax[k], c[k]); }
  }
  int axis = 0;
  R best = cmax[0] - cmin[0];
  for (int k = 1; k < D; ++k) { R e = cmax[k] - cmin[k]; if (e > best) { best = e; axis = k; } }
  if (best <= Eps<R>::v) return node;
  // :152-153: stable sort by the centroid on `axis`.
  sort_by_key(cx.keys.data() + (size_t)axis * cx.n, idx, start, count);
  const uint32_t mid = start + count / 2;  // :154

  uint32_t a, b;
  if (count >= PAR_MIN_COUNT && depth < PAR_MAX_DEPTH) {
    Local<R> la, lb;
    auto fa = std::async(std::launch::async, [&] { build_range(cx, idx, start, mid, depth + 1, la); });
    build_range(cx, idx, mid, end, depth + 1, lb);
    fa.get();
    // Splice both subtrees in pre-order, shifting their interior child links.
    auto splice = [&](Local<R>& l) {
      const uint32_t base = (uint32_t)out.ca.size();
      out.bounds.insert(out.bounds.end(), l.bounds.begin(), l.bounds.end());
      for (size_t i = 0; i < l.ca.size(); ++i) {
        bool leaf = l.pc[i] > 0;
        out.ca.push_back(leaf ? l.ca[i] : l.ca[i] + base);
        out.cb.push_back(leaf ? l.cb[i] : l.cb[i] + base);
        out.pc.push_back(l.pc[i]);
      }
      return base;
    };
    a = splice(la);
    b = splice(lb);
  } else {
    a = build_range(cx, idx, start, mid, depth + 1, out);
    b = build_range(cx, idx, mid, end, depth + 1, out);
  }
  out.ca[node] = a;
  out.cb[node] = b;
  out.pc[node] = 0;
  return node;
}

}  // namespace

int prim_stride(int dim, int kind) {
  switch (kind) {
    case 0: return 2 * dim;
    case 1: return dim + 1;
    case 2: return 9;
    default: return -1;
  }
}

int device_prim_stride(int dim, int kind) {
  switch (kind) {
    case 0: return 2 * dim;
    case 1: return dim + 1;
    case 2: return 12;
    default: return -1;
  }
}

void prim_boxes(int dim, int kind, const float* prims, uint64_t n, std::vector<float>& boxes) {
  const int D = dim;
  boxes.resize((size_t)n * 2 * D);
  const int s = prim_stride(dim, kind);
  for (uint64_t i = 0; i < n; ++i) {
    const float* p = prims + i * s;
    float* b = boxes.data() + i * 2 * D;
    if (kind == 0) {
      std::memcpy(b, p, 2 * D * sizeof(float));
    } else if (kind == 1) {
      // Aabb::new(c - splat(r), c + splat(r))
      for (int k = 0; k < D; ++k) { b[k] = p[k] - p[D]; b[D + k] = p[k] + p[D]; }
    } else {
      // union_point of v0, v1, v2 folded from EMPTY
      for (int k = 0; k < 3; ++k) {
        b[k] = fminf(fminf(fminf(INFINITY, p[k]), p[3 + k]), p[6 + k]);
        b[3 + k] = fmaxf(fmaxf(fmaxf(-INFINITY, p[k]), p[3 + k]), p[6 + k]);
      }
    }
  }
}

namespace {

template <class R> void build_tree_t(int dim, const R* boxes, uint32_t n, RefTreeT<R>& out) {
  out = RefTreeT<R>();
  out.dim = dim;
  out.prim_indices.resize(n);
  for (uint32_t i = 0; i < n; ++i) out.prim_indices[i] = i;
  if (n == 0) return;
  Ctx<R> cx;
  cx.dim = dim; cx.boxes = boxes; cx.n = n;
  cx.cent.resize((size_t)n * dim);
  cx.keys.resize((size_t)n * dim);
  for (uint32_t i = 0; i < n; ++i) {
    const R* b = boxes + (size_t)i * 2 * dim;
    for (int k = 0; k < dim; ++k) {
      R c = (b[k] + b[dim + k]) * R(0.5);   // Aabb::center, aabb.rs:56-58
      cx.cent[(size_t)i * dim + k] = c;
      cx.keys[(size_t)k * n + i] = total_ukey(c);
    }
  }
  Local<R> l;
  l.ca.reserve(2 * (size_t)n / 3 + 1);
  build_range(cx, out.prim_indices.data(), 0, n, 0, l);
  out.bounds.swap(l.bounds);
  out.child_a.swap(l.ca);
  out.child_b.swap(l.cb);
  out.prim_count.swap(l.pc);
}

}  // namespace

void build_tree(int dim, const float* boxes, uint32_t n, RefTree& out) { build_tree_t(dim, boxes, n, out); }
void build_tree(int dim, const double* boxes, uint32_t n, RefTreeT<double>& out) { build_tree_t(dim, boxes, n, out); }

template <class R> void tree_structural(const RefTreeT<R>& t, uint64_t* leaves, uint64_t* prims, uint64_t* depth_out) {
  *leaves = *prims = *depth_out = 0;
  const size_t nn = t.node_count();
  if (!nn) return;
  std::vector<uint64_t> depth(nn, 0);
  for (size_t i = nn; i-- > 0;) {
    if (t.prim_count[i] > 0) { (*leaves)++; *prims += t.prim_count[i]; depth[i] = 1; }
    else depth[i] = 1 + std::max(depth[t.child_a[i]], depth[t.child_b[i]]);
  }
  *depth_out = depth[0];
}
template void tree_structural<float>(const RefTreeT<float>&, uint64_t*, uint64_t*, uint64_t*);
template void tree_structural<double>(const RefTreeT<double>&, uint64_t*, uint64_t*, uint64_t*);

float tree_interior_surface_sum(const RefTree& t) {
  const int D = t.dim;
  float s = 0.0f;
  for (size_t i = 0; i < t.node_count(); ++i) {
    if (t.prim_count[i] > 0) continue;
    const float* b = t.bounds.data() + i * 2 * D;
    float sum = 0.0f;   // Aabb::half_surface, aabb.rs:84-97
    for (int a = 0; a < D; ++a) {
      float prod = 1.0f;
      for (int c = 0; c < D; ++c) if (c != a) prod = prod * (b[D + c] - b[c]);
      sum = sum + prod;
    }
    s = s + sum;
  }
  return s;
}

void pack_tree(const RefTree& t, Packed& out) {
  const int D = t.dim;
  const int NF = (D + 1) * 4;
  out = Packed();
  out.dim = D;
  const size_t nn = t.node_count();
  out.root.assign(16, 0.0f);
  out.slot_of_ref.assign(nn, 0xFFFFFFFFu);
  if (nn == 0) return;
  if (t.prim_indices.size() > LEAF_FIRST_MASK) throw std::runtime_error("too many primitives for the leaf encoding (max 2^27)");

  // Pair index of every interior reference node, in reference pre-order.
  std::vector<uint32_t> pair_of(nn, 0xFFFFFFFFu);
  uint32_t np = 0;
  for (size_t i = 0; i < nn; ++i) if (t.prim_count[i] == 0) pair_of[i] = np++;
  out.nodes.reserve((size_t)np * NF + 64);
  out.nodes.assign((size_t)np * NF, 0.0f);
  std::vector<uint32_t> parent(np, 0xFFFFFFFFu);
  std::vector<uint32_t> depth(np, 0);

  auto set_slot = [&](uint32_t p, int slot, const float* b, uint32_t ref) {
    float* nd = out.nodes.data() + (size_t)p * NF;
    for (int k = 0; k < D; ++k) { nd[4 * k + 2 * slot] = b[k]; nd[4 * k + 2 * slot + 1] = b[D + k]; }
    std::memcpy(&nd[4 * D + slot], &ref, 4);
  };
  auto new_pair = [&](uint32_t par, uint32_t dep) {
    out.nodes.resize(out.nodes.size() + NF, 0.0f);
    parent.push_back(par);
    depth.push_back(dep);
    return np++;
  };
  // Synthetic pair nodes (leaf chains) carry 1 in the last lane.
  std::vector<uint32_t> synth;
  synth.assign(np, 0u);
  // The leaf ref for a reference leaf node `li` placed under pair `par`.
  std::function<uint32_t(uint32_t, uint32_t, uint32_t, uint32_t, const float*)> leaf_chain;
  leaf_chain = [&](uint32_t first, uint32_t count, uint32_t par, uint32_t dep, const float* box) -> uint32_t {
    if (count <= LEAF_MAX_COUNT) return leaf_ref(first, count);
    // Synthetic pair: slot a = first LEAF_MAX_COUNT prims, slot b = the rest.
    // Both carry the whole leaf's box: valid (it contains each part) and exact
    // after a device refit, which recomputes parts from their primitives.
    uint32_t p = new_pair(par, dep);
    synth.push_back(1u);
    uint32_t ra = leaf_ref(first, LEAF_MAX_COUNT);
    uint32_t rb = leaf_chain(first + LEAF_MAX_COUNT, count - LEAF_MAX_COUNT, p, dep + 1, box);
    set_slot(p, 0, box, ra);
    set_slot(p, 1, box, rb);
    return p;
  };

  // Depth of pair nodes from the reference parent links (pre-order: parents first).
  for (size_t i = 0; i < nn; ++i) {
    if (t.prim_count[i] != 0) continue;
    uint32_t p = pair_of[i];
    uint32_t kids[2] = {t.child_a[i], t.child_b[i]};
    for (int s = 0; s < 2; ++s) {
      uint32_t c = kids[s];
      const float* cbx = t.bounds.data() + (size_t)c * 2 * D;
      uint32_t ref;
      if (t.prim_count[c] == 0) {
        ref = pair_of[c];
        parent[ref] = p;
        depth[ref] = depth[p] + 1;
      } else {
        ref = leaf_chain(t.child_a[c], t.prim_count[c], p, depth[p] + 1, cbx);
      }
      set_slot(p, s, cbx, ref);
      out.slot_of_ref[c] = p * 2 + s;
    }
  }
  uint32_t root_ref;
  if (t.prim_count[0] == 0) root_ref = 0;
  else root_ref = leaf_chain(t.child_a[0], t.prim_count[0], 0xFFFFFFFFu, 0, t.bounds.data());
  for (int k = 0; k < 2 * D; ++k) out.root[k] = t.bounds[k];
  std::memcpy(&out.root[2 * D], &root_ref, 4);
  // parent links (for bottom-up refit) in the spare lane of each pair node
  for (uint32_t p = 0; p < np; ++p) {
    std::memcpy(&out.nodes[(size_t)p * NF + 4 * D + 2], &parent[p], 4);
    std::memcpy(&out.nodes[(size_t)p * NF + 4 * D + 3], &synth[p], 4);
  }
  out.n_pairs = np;
  out.n_synth = 0;
  for (uint32_t p = 0; p < np; ++p) out.n_synth += synth[p];

  // Level lists, deepest first, for the level-synchronous device refit.
  uint32_t maxd = 0;
  for (uint32_t p = 0; p < np; ++p) maxd = std::max(maxd, depth[p]);
  std::vector<std::vector<uint32_t>> lv(np ? maxd + 1 : 0);
  for (uint32_t p = 0; p < np; ++p) lv[depth[p]].push_back(p);
  out.level_off.push_back(0);
  for (size_t d = lv.size(); d-- > 0;) {
    out.levels.insert(out.levels.end(), lv[d].begin(), lv[d].end());
    out.level_off.push_back((uint32_t)out.levels.size());
  }
  // A near-first walk pushes at most one entry per pair level.
  out.max_stack = np ? maxd + 2 : 1;
}

void pack_prims(int dim, int kind, const float* prims, const RefTree& t, std::vector<float>& out) {
  const int s = prim_stride(dim, kind);
  const int ds = device_prim_stride(dim, kind);
  const size_t n = t.prim_indices.size();
  out.assign(n * ds, 0.0f);
  for (size_t k = 0; k < n; ++k) {
    const uint32_t id = t.prim_indices[k];
    const float* p = prims + (size_t)id * s;
    float* q = out.data() + k * ds;
    if (kind == 2) {
      q[0] = p[0]; q[1] = p[1]; q[2] = p[2]; std::memcpy(&q[3], &id, 4);
      q[4] = p[3]; q[5] = p[4]; q[6] = p[5]; q[7] = 0.0f;
      q[8] = p[6]; q[9] = p[7]; q[10] = p[8]; q[11] = 0.0f;
    } else {
      std::memcpy(q, p, s * sizeof(float));
    }
  }
}

}  // namespace sqh
