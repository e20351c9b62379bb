// host_tree.h -- host side of the engine: the reference median-split build and
// the re-pack of its node array into the device layout.
#pragma once

#include <cstdint>
#include <vector>

namespace sqh {

constexpr int MAX_DIM = 8;

// A tree in the reference layout: Bvh::nodes() / Bvh::prim_indices()
// (crates/spatial-query/src/accel/tree.rs:20-26, bvh.rs:465-475), over the
// reference's Scalar: float, or double under its `f64` feature
// (crates/spatial-query/src/maths/point.rs:14-18).
template <class R> struct RefTreeT {
  int dim = 0;
  std::vector<R> bounds;            // [nn][2*dim] min then max
  std::vector<uint32_t> child_a, child_b, prim_count;
  std::vector<uint32_t> prim_indices;
  size_t node_count() const { return child_a.size(); }
};
using RefTree = RefTreeT<float>;

// Primitive boxes (QueryGeometry::world_aabb) of the built-in kinds, [n][2*dim].
void prim_boxes(int dim, int kind, const float* prims, uint64_t n, std::vector<float>& boxes);
int prim_stride(int dim, int kind);  // floats per primitive in the caller's layout

// tree::build (tree.rs:39-49, 112-163): same topology node for node.
void build_tree(int dim, const float* boxes, uint32_t n, RefTree& out);
void build_tree(int dim, const double* boxes, uint32_t n, RefTreeT<double>& out);

// TreeStats pieces (tree.rs:81-110).
template <class R> void tree_structural(const RefTreeT<R>& t, uint64_t* leaves, uint64_t* prims, uint64_t* depth);
float tree_interior_surface_sum(const RefTree& t);

// Device layout.
//
// Interior nodes become "pair nodes" holding both children's boxes, so one
// fetch of (dim+1)*16 bytes tests both children:
//   float4 axis[k] = {a.min[k], a.max[k], b.min[k], b.max[k]}   k < dim
//   uint4  refs    = {ref_a, ref_b, parent, synthetic}
// A child ref is either a pair-node index (bit 31 clear) or a leaf
// 0x80000000 | (count-1) << 27 | first, covering primitives
// [first, first+count) of the leaf-ordered arrays (count <= 16; larger leaves
// are split into chains of synthetic pair nodes).
constexpr uint32_t LEAF_BIT = 0x80000000u;
constexpr uint32_t LEAF_MAX_COUNT = 16;
constexpr uint32_t LEAF_FIRST_MASK = (1u << 27) - 1;
inline uint32_t leaf_ref(uint32_t first, uint32_t count) { return LEAF_BIT | ((count - 1) << 27) | first; }

struct Packed {
  int dim = 0;
  std::vector<float> nodes;          // [np][(dim+1)*4]
  uint32_t n_pairs = 0;
  std::vector<float> root;           // [2*dim] bounds, then ref bits (as float), padded to 16 floats
  std::vector<uint32_t> levels;      // pair indices sorted by depth, deepest level first
  std::vector<uint32_t> level_off;   // offsets into levels, one range per depth (deepest first)
  // For mapping device boxes back to reference node order: for every
  // reference node except the root, the (pair, slot) holding its box.
  std::vector<uint32_t> slot_of_ref; // pair*2 + slot; root = 0xFFFFFFFF
  uint32_t max_stack = 0;            // pair-tree depth bound for the traversal stack
  uint32_t n_synth = 0;              // synthetic pair nodes (chains of over-full leaves)
};

// Re-pack a reference tree for the device.
void pack_tree(const RefTree& t, Packed& out);

// Leaf-ordered primitive arrays for the device, from caller-layout primitives:
// AABB: [min D][max D]; SPHERE: [c D][r]; TRIANGLE: v0.xyz,id, v1.xyz,0, v2.xyz,0.
int device_prim_stride(int dim, int kind);
void pack_prims(int dim, int kind, const float* prims, const RefTree& t, std::vector<float>& out);

}  // namespace sqh
