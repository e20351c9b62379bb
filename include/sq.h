/*
 * sq.h -- C ABI of the MI355X (gfx950) batched BVH query engine.
 *
 * This is the drop-in boundary for the batched query path of
 * grimandgreedy/spatial-query.  The reference binds a device backend at two
 * places, and every entry point below replaces one of them:
 *
 *   - the device seam `GpuRaycaster` (crates/spatial-query/src/gpu/mod.rs:252-463):
 *     new / headless / refit / raycast_nearest_batch, fed by Bvh::nodes() and
 *     Bvh::prim_indices() (crates/spatial-query/src/accel/bvh.rs:465-475);
 *   - the plugin trait `QueryBackend<D, G>` (crates/spatial-query/src/dispatch.rs:217-237),
 *     whose raycast_nearest_batch answers one Option<Hit> per ray in order.
 *
 * Conventions
 *   - Every function returns SQ_OK (0) or an SQ_ERR_* code; sq_last_error()
 *     gives the message of the last failure on the calling thread.
 *   - Rays are `float[n][2*D]`: origin then direction (Ray<D>, ray.rs:9-13).
 *     Directions are used as given (Ray::new_unnormalized); callers that want
 *     world-unit toi pass unit directions, as the reference does.
 *   - Query input and output pointers are DEVICE pointers (hipMalloc'd or
 *     torch tensors).  Queries are enqueued on the context's stream and return
 *     without waiting; sq_ctx_synchronize() waits.  Calls on one context are
 *     serialised by the caller; separate contexts are independent.
 *   - A miss writes leaf -1 and toi +inf (the reference's `None`).
 *   - Leaf ids are the provider's leaf indices 0..n-1 (QueryGeometry::id is
 *     the leaf index for the built-in kinds).
 *   - All arithmetic is FP32 with the reference's operation order (FP64 for
 *     the sq_bvh64 entries, the reference's `f64` feature); results are
 *     bit-identical to the reference CPU path for finite inputs.
 */
#ifndef SQ_H
#define SQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SQ_OK 0
#define SQ_ERR_INVALID 1      /* bad argument                                   */
#define SQ_ERR_NO_DEVICE 2    /* no usable GPU (the reference's `headless` None) */
#define SQ_ERR_HIP 3          /* a HIP runtime call failed                      */
#define SQ_ERR_UNSUPPORTED 4  /* dimension / kind combination not instantiated  */
#define SQ_ERR_CAPACITY 5     /* variable-length output larger than capacity    */

/* Built-in providers (the QueryGeometry narrow phases, geometry.rs:19-98).
 * Primitive layouts (host or device, leaf-indexed):
 *   SQ_PRIM_AABB     float[n][2*D]  min then max     toi = slab entry (aabb.rs:118-148)
 *   SQ_PRIM_SPHERE   float[n][D+1]  centre, radius   differential.rs:61-82
 *   SQ_PRIM_TRIANGLE float[n][9]    v0, v1, v2 (D=3) Moeller-Trumbore, viewport-lib-query/src/lib.rs:279-306
 * SPHERE and TRIANGLE hits are kept only when their toi lies inside the slab
 * interval of the primitive's own AABB, which makes every result independent
 * of traversal order (and equal to the naive scan). */
#define SQ_PRIM_AABB 0
#define SQ_PRIM_SPHERE 1
#define SQ_PRIM_TRIANGLE 2

typedef struct sq_ctx sq_ctx;
typedef struct sq_bvh sq_bvh;
typedef struct sq_tlas sq_tlas;

/* QueryFilter (crates/spatial-query/src/query/filter.rs:12-29).  The built-in
 * kinds accept a leaf when (layers[leaf] & layer_mask) != 0 and leaf != exclude.
 * exclude = -1 is `None`.  The default filter is {~0ull, -1}. */
typedef struct sq_filter {
  uint64_t layer_mask;
  int64_t exclude;
} sq_filter;

/* TreeStats (crates/spatial-query/src/accel/stats.rs:15-32). */
typedef struct sq_tree_stats {
  uint64_t node_count, leaf_count, prim_count, max_depth, bytes;
  float quality;
} sq_tree_stats;

const char* sq_last_error(void);
int sq_version(void);

/* ---- context: GpuRaycaster::headless / new (gpu/mod.rs:252-324) ---------- */
int sq_ctx_create(int device, sq_ctx** out);   /* SQ_ERR_NO_DEVICE when absent */
/* Releases the caller's handle.  Trees built on the context keep it alive
 * until they are destroyed too, in any order. */
int sq_ctx_destroy(sq_ctx* ctx);
/* Enqueue all further work on `stream` (a hipStream_t of the same device, e.g.
 * the caller's framework stream; NULL is HIP's legacy default stream) instead
 * of the context's own stream.  sq_ctx_reset_stream goes back to its own.
 * A context's calls share its scratch buffers and pool counters: work it
 * enqueued on one stream must be finished, or ordered before, when work is
 * enqueued on another (the calls of one stream are ordered by the stream). */
int sq_ctx_set_stream(sq_ctx* ctx, void* stream);
int sq_ctx_reset_stream(sq_ctx* ctx);
int sq_ctx_synchronize(sq_ctx* ctx);
/* The hipStream_t all work of this context is enqueued on. */
void* sq_ctx_stream(sq_ctx* ctx);
/* Which hits the context's queries return (the reference has one behaviour,
 * SQ_HITS_REFERENCE; the other is this engine's opt-in).
 *   SQ_HITS_REFERENCE (default)  every query returns what the reference's CPU
 *       path returns for the built-in provider: the narrow phases exactly as
 *       the reference's providers (no extra test), and nearest-hit rays and
 *       shape casts over leaves whose hits may fall outside their own AABB
 *       (SPHERE, TRIANGLE, consumer narrow phases, every instance kind) walk
 *       the reference's order (bvh.rs:73-124: push child a, then b, pop LIFO,
 *       skip what lies beyond the current best).  AABB leaves walk near-first
 *       in both modes, which gives the same hits.
 *   SQ_HITS_GUARDED  SPHERE and TRIANGLE hits (and instance hits, against the
 *       instance's world box) count only when their toi lies inside the slab
 *       interval of their own AABB; then any walk order gives one answer, and
 *       nearest queries walk near-first (fewer nodes).  It differs from the
 *       reference only on rays whose reference answer lies outside its box
 *       (40 of the 8,388,608 rays of the 2M-triangle heightfield workload).
 *       The learned proposal's seeded walks (sq_raycast_nearest_proposed,
 *       sq_tlas_raycast_nearest_proposed) always answer in this mode.
 * All-hits, crossings, overlap and cast-all lists hold the reference's hits in
 * both modes (their walks never shrink a limit, so order does not matter). */
#define SQ_HITS_REFERENCE 0
#define SQ_HITS_GUARDED 1
int sq_ctx_set_hit_mode(sq_ctx* ctx, int mode);
int sq_ctx_get_hit_mode(const sq_ctx* ctx);   /* -1 for a NULL ctx */

/* ---- host tree: Bvh (accel/bvh.rs), canonical, no device ------------------ */
/* Bvh::build(&G) (bvh.rs:31-37, tree.rs:39-163) for a built-in provider: the
 * reference median-split build on the host, same topology node for node.
 * prims is a HOST pointer in the kind's layout.  Needs no GPU. */
typedef struct sq_tree sq_tree;
int sq_tree_build(int dim, int kind, const float* prims, uint64_t n_prims, sq_tree** out);
/* A tree the caller already built in the reference layout (Bvh::nodes():
 * bounds float[nn][2D], child_a, child_b, prim_count; Bvh::prim_indices()),
 * validated and copied.  Host pointers. */
int sq_tree_from_nodes(int dim, const float* node_bounds, const uint32_t* child_a, const uint32_t* child_b,
                       const uint32_t* prim_count, uint64_t n_nodes, const uint32_t* prim_indices,
                       uint64_t n_prims, sq_tree** out);
/* Bvh::refit(&G) on the host (tree.rs:54-75): new primitives, same count. */
int sq_tree_refit(sq_tree* tree, int kind, const float* prims);
int sq_tree_get_stats(const sq_tree* tree, sq_tree_stats* out);   /* Bvh::stats, bvh.rs:478-489 */
/* Bvh::nodes() / prim_indices() (bvh.rs:465-475).  Any pointer may be NULL. */
int sq_tree_export(const sq_tree* tree, float* node_bounds, uint32_t* child_a, uint32_t* child_b,
                   uint32_t* prim_count, uint32_t* prim_indices);
int sq_tree_destroy(sq_tree* tree);

/* ---- device tree: GpuRaycaster (gpu/mod.rs:252-341) ----------------------- */
/* GpuRaycaster::new(device, queue, &Bvh, &G): re-pack the host tree into the
 * device layout and upload it with the primitives (HOST pointer, kind layout).
 * The host tree is copied; the caller keeps ownership of `tree`. */
int sq_bvh_upload(sq_ctx* ctx, const sq_tree* tree, int kind, const float* prims, sq_bvh** out);
/* sq_tree_build + sq_bvh_upload in one call. */
int sq_bvh_build(sq_ctx* ctx, int dim, int kind, const float* prims, uint64_t n_prims, sq_bvh** out);
int sq_bvh_destroy(sq_bvh* bvh);
/* Per-leaf layer bits (host pointer, n_prims entries, or NULL for all-ones). */
int sq_bvh_set_layers(sq_bvh* bvh, const uint64_t* layers);
int sq_bvh_get_stats(const sq_bvh* bvh, sq_tree_stats* out);   /* Bvh::stats of the uploaded tree */
/* Node bounds as currently held ON THE DEVICE, mapped back to reference node
 * order (float[nn][2D], host pointer; synchronises).  Shows a device refit. */
int sq_bvh_download_bounds(sq_bvh* bvh, float* node_bounds);
/* Bvh::nodes() / prim_indices() (bvh.rs:465-475) of a device tree: node
 * bounds as sq_bvh_download_bounds (the current device boxes), child_a,
 * child_b, prim_count (node_count entries each) and prim_indices (n_prims).
 * Any pointer may be NULL. */
int sq_bvh_export(sq_bvh* bvh, float* node_bounds, uint32_t* child_a, uint32_t* child_b, uint32_t* prim_count,
                  uint32_t* prim_indices);
/* GpuRaycaster::refit / Bvh::refit(&G) (gpu/mod.rs:326-341, tree.rs:54-75):
 * new primitives in the kind's layout, DEVICE pointer, same leaf count;
 * topology kept; bottom-up refit on the device, bit-identical to the host
 * refit.  Asynchronous. */
int sq_bvh_refit(sq_bvh* bvh, const float* prims_device);
/* Bvh::build (bvh.rs:31-37, tree.rs:39-49,112-163) from primitives already in
 * device memory, built on the device: the same tree node for node as
 * sq_tree_build + sq_bvh_upload (same topology, prim_indices, bounds and
 * stats), level by level with one radix sort per level.  Synchronises.  Not
 * for graph capture. */
int sq_bvh_build_device(sq_ctx* ctx, int dim, int kind, const float* prims_device, uint64_t n_prims, sq_bvh** out);
/* Rebuild in place from moved primitives (device memory, the same count and
 * kind): Bvh::build again on the new geometry, where sq_bvh_refit keeps the
 * topology (tree.rs:54-75) and lets the tree loosen as primitives move.  The
 * device arrays keep their addresses (a device-built tree reserves room for
 * any tree of its primitive count), so batch sets and captured graphs over
 * the tree stay valid; a batch set sized its traversal stack when it was
 * made, so make it again if the rebuilt tree is deeper (stats max_depth). */
int sq_bvh_rebuild(sq_bvh* bvh, const float* prims_device);
/* Interior surface sum (tree.rs:104-110, the Tlas::needs_rebuild figure,
 * tlas.rs:56-60) of the tree's current device bounds, computed on the device:
 * each node's half surface in f32 as the reference computes it, summed in f64
 * in a fixed order (deterministic; within f32 rounding of the reference's
 * running f32 sum).  Synchronises. */
int sq_bvh_surface_sum(sq_bvh* bvh, double* out);
/* Refit, then rebuild when the refit tree has loosened: the policy of
 * Tlas::needs_rebuild (tlas.rs:56-60) applied to a moving single-level tree.
 * sq_bvh_refit on prims_device, then, when the device surface sum exceeds
 * factor (>= 1) times the sum of the tree as last built (taken at the first
 * call, and after every rebuild), sq_bvh_rebuild on the same primitives;
 * *rebuilt = 1 then.  Synchronises (not for graph capture); a rebuild keeps
 * the device arrays of a device-built tree, and sq_batch_check says whether
 * a batch set (or a graph of its submits) over the tree is still valid. */
int sq_bvh_refit_auto(sq_bvh* bvh, const float* prims_device, float factor, int* rebuilt);

/* ------------------------------------------------------------------------
 * Consumer narrow phases: GpuGeometry::narrow_wgsl (gpu/mod.rs:86-99,261),
 * where the reference splices a consumer's WGSL narrow test into its traversal
 * shader.  `source` is HIP C++ defining
 *   __device__ bool sq_narrow(const float* prim, const float* origin,
 *                             const float* dir, float max_toi, float* toi);
 * (QueryGeometry::test_ray, geometry.rs:33-43: true with *toi in [0, max_toi]
 * on a hit), compiled at run time (hipRTC, gfx950) together with the engine's
 * nearest-hit walk and its flags (no contraction, f32 denormals; plain `/` and
 * __builtin_sqrtf are correctly rounded).  No <cstring> or <cmath> there:
 * use builtins.  SQ_ERR_INVALID with the compiler log in sq_last_error() when
 * it does not compile.  A hit counts only when its toi lies in the slab
 * interval of the primitive's own world box (the engine's order-independence
 * guard: a consumer whose hits lie in its boxes loses none).
 * The source may also define every crossing of a ray with a primitive
 * (QueryGeometry::test_ray_crossings, geometry.rs:45-63):
 *   #define SQ_NARROW_CROSSINGS 1
 *   template <class Out> __device__ void sq_narrow_crossings(const float* prim,
 *       const float* origin, const float* dir, float max_toi, Out& out);
 * calling out(toi) once per crossing in any order (crossings outside the box
 * interval are dropped by the same guard); without it a crossing query gets
 * the sq_narrow hit, the reference's default.  Likewise optional: shape casts
 * (QueryGeometry::test_shape_cast, geometry.rs:65-77)
 *   #define SQ_NARROW_CAST 1
 *   __device__ bool sq_narrow_cast(const float* prim, const float* origin,
 *       const float* dir, const float* probe_min, const float* probe_max,
 *       float max_toi, float* toi);
 * (a contact counts inside the probe-grown slab interval of the primitive's
 * box; without it no cast finds a contact, the trait's default) and overlap
 * (test_overlap, geometry.rs:79-87)
 *   #define SQ_NARROW_OVERLAP 1
 *   __device__ bool sq_narrow_overlap(const float* prim, const float* region_min,
 *       const float* region_max);
 * asked for primitives whose world box meets the region (without it, that box
 * test alone, the trait's default). */
typedef struct sq_narrow sq_narrow;
int sq_narrow_create(sq_ctx* ctx, int dim, const char* source, uint32_t prim_floats, sq_narrow** out);
int sq_narrow_destroy(sq_narrow* narrow);
/* The tree over n consumer primitives (host memory): world boxes [n][2*dim]
 * (QueryGeometry::world_aabb) drive the reference median split; prims
 * [n][prim_floats] are what sq_narrow sees.  The tree answers only
 * sq_raycast_nearest_consumer and sq_raycast_all_consumer with its narrow
 * phase. */
int sq_bvh_build_consumer(sq_narrow* narrow, const float* boxes, const float* prims, uint64_t n, sq_bvh** out);
/* Nearest hit per ray through the consumer's narrow phase (raycast_nearest_batch,
 * bvh.rs:134-193, with the consumer's test_ray); device buffers as
 * sq_raycast_nearest: rays [n][2*dim], leaf (-1 on a miss), toi (+inf). */
int sq_raycast_nearest_consumer(sq_narrow* narrow, sq_bvh* bvh, const float* rays, uint64_t n, float max_toi,
                                int32_t* out_leaf, float* out_toi);
/* Bvh::raycast_all (bvh.rs:196-259) with the consumer's test_ray, or
 * raycast_crossings (bvh.rs:264-328) with its test_ray_crossings when
 * crossings != 0: CSR results, capacity and total as sq_raycast_all. */
int sq_raycast_all_consumer(sq_narrow* narrow, sq_bvh* bvh, const float* rays, uint64_t n, float max_toi, int crossings,
                            uint64_t* offsets, int32_t* out_leaf, float* out_toi, uint64_t capacity, uint64_t* total);
/* Bvh::shapecast_nearest / shapecast_all (bvh.rs:331-414) through the
 * consumer's sq_narrow_cast: casts float[n][4*dim] as sq_shapecast_nearest;
 * nearest leaf -1 / toi +inf on a miss; all = CSR as sq_raycast_all. */
int sq_shapecast_nearest_consumer(sq_narrow* narrow, sq_bvh* bvh, const float* casts, uint64_t n, float max_toi,
                                  int32_t* out_leaf, float* out_toi);
int sq_shapecast_all_consumer(sq_narrow* narrow, sq_bvh* bvh, const float* casts, uint64_t n, float max_toi,
                              uint64_t* offsets, int32_t* out_leaf, float* out_toi, uint64_t capacity, uint64_t* total);
/* Bvh::overlap (bvh.rs:418-456) through the consumer's sq_narrow_overlap:
 * CSR of ascending leaf ids as sq_overlap. */
int sq_overlap_consumer(sq_narrow* narrow, sq_bvh* bvh, const float* regions, uint64_t n, uint64_t* offsets,
                        int32_t* out_leaf, uint64_t capacity, uint64_t* total);

/* ---- ray queries (bvh.rs:51-328) ------------------------------------------ */
/* max_toi_per_ray (device, n floats) overrides max_toi when not NULL.
 * filter may be NULL (default filter). */

/* Bvh::raycast_nearest over a batch (bvh.rs:51-124, :134-193).  out_leaf and
 * out_toi are required; out_point, out_normal (float[n][D]) and out_sub
 * (sub-object: the leaf for SPHERE/TRIANGLE, -1 for AABB) may be NULL. */
int sq_raycast_nearest(sq_bvh* bvh, const float* rays, uint64_t n, float max_toi, const float* max_toi_per_ray,
                       const sq_filter* filter, int32_t* out_leaf, float* out_toi, float* out_point,
                       float* out_normal, int32_t* out_sub);
/* sq_raycast_nearest without allocating (no counterpart in the reference):
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as sq_raycast_nearest; the walk's counters and stack
 * spill live in a file-scope device array of SQ_NOALLOC_SCRATCH_BYTES on each
 * device (held until the process ends) instead of the context's buffers, and
 * the call creates no timing events.  Limits: the tree's stack must fit the
 * array (the LDS part of the stack grows to 64 entries first, then the grid
 * shrinks; SQ_ERR_UNSUPPORTED when even one block does not fit), and the pooled
 * walk always runs (rays_per_lane 0, ray_order and the 4-wide walk are not used).
 * Calls that overlap on one device race on the array: order them (one stream,
 * or events).  The context must come from sq_ctx_create (which looks up the
 * array's address on its device). */
#define SQ_NOALLOC_SCRATCH_BYTES (16u << 20)
int sq_raycast_nearest_noalloc(sq_bvh* bvh, const float* rays, uint64_t n, float max_toi, const float* max_toi_per_ray,
                               const sq_filter* filter, int32_t* out_leaf, float* out_toi, float* out_point,
                               float* out_normal, int32_t* out_sub);
/* Bvh::raycast_nearest_stats / raycast_all_stats (bvh.rs:61-70, 206-215,
 * QueryStats stats.rs:39-50): per ray, out_stats[i] = {nodes_visited,
 * aabb_tests, narrow_tests, hits} as the reference counts them on its own walk
 * (every popped node a visit and an AABB test, every filter-accepted primitive
 * of a passing leaf a narrow test, every narrow test that hits a hit), from a
 * walk in the reference's order over the reference's hits whatever the
 * context's hit mode; device uint32[n][4].  The nearest form also writes the
 * nearest hit (leaf, toi); the all-hits form only the counters (the hit lists
 * come from sq_raycast_all).  Built-in kinds (AABB, SPHERE, TRIANGLE).  A
 * diagnostic walk, one ray per thread: not the fast path. */
int sq_raycast_nearest_stats(sq_bvh* bvh, const float* rays, uint64_t n, float max_toi, const float* max_toi_per_ray,
                             const sq_filter* filter, int32_t* out_leaf, float* out_toi, uint32_t* out_stats);
int sq_raycast_all_stats(sq_bvh* bvh, const float* rays, uint64_t n, float max_toi, const sq_filter* filter,
                         uint32_t* out_stats);
/* Occlusion: out_hit[i] = raycast_nearest(..).is_some() (the reference's any-hit
 * idiom, examples-viewport/examples/many_rays.rs:193-201), ending each ray at
 * its first accepted hit. */
int sq_raycast_any(sq_bvh* bvh, const float* rays, uint64_t n, float max_toi, const float* max_toi_per_ray,
                   const sq_filter* filter, uint8_t* out_hit);
/* Bvh::raycast_all (bvh.rs:196-259), or raycast_crossings (bvh.rs:264-328)
 * when crossings != 0.  Results in CSR form: offsets (device, n+1, written),
 * then out_leaf/out_toi (device, capacity entries), each ray's hits sorted by
 * (toi total order, leaf) as hit_ordering (query/hit.rs:77-84).  *total (host)
 * receives the hit count; if it exceeds capacity nothing is filled and
 * SQ_ERR_CAPACITY is returned.  Synchronises once (to read the total). */
int sq_raycast_all(sq_bvh* bvh, const float* rays, uint64_t n, float max_toi, const sq_filter* filter,
                   int crossings, uint64_t* offsets, int32_t* out_leaf, float* out_toi, uint64_t capacity,
                   uint64_t* total);
/* Bvh::overlap (bvh.rs:418-456): regions float[n][2D] (min then max); each
 * query's leaves in ascending order, CSR as above. */
int sq_overlap(sq_bvh* bvh, const float* regions, uint64_t n, const sq_filter* filter, uint64_t* offsets,
               int32_t* out_leaf, uint64_t capacity, uint64_t* total);
/* ShapeCast queries (bvh.rs:331-414, query/shapecast.rs:26-73): casts are
 * float[n][4D] = origin, dir, probe box min, probe box max (offsets from the
 * reference point).  AABB leaves answer with the slab of the grown box;
 * SPHERE leaves treat the probe as a ball of radius probe_max[0]; TRIANGLE
 * leaves keep the reference default (no contact). */
int sq_shapecast_nearest(sq_bvh* bvh, const float* casts, uint64_t n, float max_toi, const float* max_toi_per_cast,
                         const sq_filter* filter, int32_t* out_leaf, float* out_toi, float* out_point);
int sq_shapecast_all(sq_bvh* bvh, const float* casts, uint64_t n, float max_toi, const float* max_toi_per_cast,
                     const sq_filter* filter, uint64_t* offsets, int32_t* out_leaf, float* out_toi,
                     uint64_t capacity, uint64_t* total);

/* Order-independent minimal work of each ray at limit t_limit[i] (the final
 * nearest toi, or max_toi on a miss): out_nodes = node boxes that must be
 * tested, out_prims = primitives that must be tested.  Feeds the algorithmic
 * bytes of the roofline.  The reference's QueryStats (accel/stats.rs:39-50)
 * counts the same events in its own traversal order. */
int sq_required_work(sq_bvh* bvh, const float* rays, uint64_t n, const float* t_limit, uint32_t* out_nodes,
                     uint32_t* out_prims);

/* ---- two-level tree: Tlas (accel/tlas.rs) -------------------------------- */
/* Built-in instance kinds (InstancedGeometry, query/instance.rs:19-106):
 *   SQ_INST_BOX   local box [-half, half] (any D), the reference's
 *                 ray_local_box provider (tests/two_level.rs:70-113);
 *   SQ_INST_MESH  a uniformly scaled triangle mesh (D = 3): its local geometry
 *                 is the mesh scaled by `scale`; the local ray is taken into the
 *                 unscaled mesh frame by o/scale, d/scale (toi unchanged) and
 *                 answered by the mesh's own device BVH (a TRIANGLE sq_bvh);
 *   SQ_INST_BALL  a ball of radius r at the local origin (D = 3), the
 *                 reference's BallInstances (tests/shape_overlap.rs:288-345):
 *                 rays by ray_ball, shape casts by the ray against the radius
 *                 sum (probe radius = the cast box's max x), overlap exact
 *                 (distance from the centre to the region <= r).
 * Boxes and meshes keep InstancedGeometry's defaults for the rest: no shape
 * cast (test_shape_cast = None), AABB-level overlap (test_overlap = true),
 * and every kind reports one crossing per instance (test_ray_crossings_local's
 * default), so Tlas::raycast_crossings is sq_tlas_raycast_all.
 * Under SQ_HITS_GUARDED an instance hit is kept only when its toi lies in the
 * slab interval of the instance's world AABB (order independence, as for the
 * primitive kinds); a cast contact only inside the probe-grown world AABB's
 * slab interval.  SQ_HITS_REFERENCE keeps what the instance test returns. */
#define SQ_INST_BOX 0
#define SQ_INST_MESH 1
#define SQ_INST_BALL 2
/* Tlas::build (tlas.rs:35-45): rot float[n][D][D] row-major (orthonormal),
 * trans float[n][D], halves float[n][D] (BOX) or radii float[n] (BALL), meshes[n_meshes] (MESH, borrowed:
 * keep them alive), mesh_of[n], scale[n].  Host pointers. */
int sq_tlas_build(sq_ctx* ctx, int dim, int inst_kind, uint64_t n_inst, const float* rot, const float* trans,
                  const float* halves, sq_bvh* const* meshes, uint32_t n_meshes, const uint32_t* mesh_of,
                  const float* scale, sq_tlas** out);
int sq_tlas_destroy(sq_tlas* tlas);
/* Tlas::refit (tlas.rs:49-54): new transforms (and halves for BOX, or NULL),
 * host pointers; world AABBs recomputed, tree refit on the device. */
int sq_tlas_refit(sq_tlas* tlas, const float* rot, const float* trans, const float* halves);
/* Tlas::needs_rebuild (tlas.rs:58-60).  Synchronises. */
int sq_tlas_needs_rebuild(sq_tlas* tlas, float factor, int* out);
int sq_tlas_get_stats(const sq_tlas* tlas, sq_tree_stats* out);   /* Tlas::stats, tlas.rs:68-79 */
int sq_tlas_download_bounds(sq_tlas* tlas, float* node_bounds);
/* Tlas::raycast_nearest (tlas.rs:82-154): out_inst = instance index, out_sub =
 * triangle index within the mesh (MESH) or -1 (BOX); normals lifted by R. */
int sq_tlas_raycast_nearest(sq_tlas* tlas, const float* rays, uint64_t n, float max_toi, const float* max_toi_per_ray,
                            int32_t* out_inst, float* out_toi, float* out_point, float* out_normal, int32_t* out_sub);
int sq_tlas_raycast_any(sq_tlas* tlas, const float* rays, uint64_t n, float max_toi, const float* max_toi_per_ray,
                        uint8_t* out_hit);
/* Tlas::raycast_all (tlas.rs:158-215), CSR as sq_raycast_all. */
int sq_tlas_raycast_all(sq_tlas* tlas, const float* rays, uint64_t n, float max_toi, uint64_t* offsets,
                        int32_t* out_inst, float* out_toi, uint64_t capacity, uint64_t* total);
/* Tlas::overlap (tlas.rs:393-432): every instance of a touched leaf node that
 * passes its test_overlap (exact for BALL, true otherwise), CSR, instances in
 * ascending order. */
int sq_tlas_overlap(sq_tlas* tlas, const float* regions, uint64_t n, uint64_t* offsets, int32_t* out_inst,
                    uint64_t capacity, uint64_t* total);
/* Tlas::shapecast_nearest (tlas.rs:306-351): casts float[n][4*D] as
 * sq_shapecast_nearest (origin, dir, probe box min, max); out_inst -1 and toi
 * +inf on a miss (every cast misses box and mesh instances); out_point
 * (optional) = origin + dir * toi. */
int sq_tlas_shapecast_nearest(sq_tlas* tlas, const float* casts, uint64_t n, float max_toi, int32_t* out_inst,
                              float* out_toi, float* out_point);
/* Tlas::shapecast_all (tlas.rs:353-391), CSR sorted by (toi, instance). */
int sq_tlas_shapecast_all(sq_tlas* tlas, const float* casts, uint64_t n, float max_toi, uint64_t* offsets,
                          int32_t* out_inst, float* out_toi, uint64_t capacity, uint64_t* total);

/* Kernel tuning knobs.  Results never depend on them (the tests check this);
 * only speed does.  The autotuner (sq_amd.autotune) picks them per scene.
 *   block          threads per block, a multiple of 64 (pooled kernels: <= 256)
 *   rays_per_lane  nearest / any / shape-cast: queries are handed out in chunks
 *                  of 64 * rays_per_lane (default 16); a grid of the waves the
 *                  chip holds resident walks them, each wave its own chunk
 *                  first, then chunks from a shared counter, refilling finished
 *                  lanes as it goes; small calls halve the chunk until every
 *                  resident wave has one; 0 selects one query per thread with
 *                  the whole stack in LDS
 *   lds_stack      entries of each lane's traversal stack kept in LDS (power of
 *                  two); deeper entries spill to device memory
 *   ray_order      1: pooled nearest / any-hit calls of >= 4096 rays walk the
 *                  rays in Morton order (a 30-bit key interleaving those
 *                  components of origin and direction that vary in the batch,
 *                  each quantized within the batch's own bounds; sorted on the
 *                  device); results still land at each ray's own index.
 *                  Not for graph capture (the sort is a library call).
 *   descend_min    pooled walks: a wave stops waiting on descents once fewer
 *                  than this many lanes (0..64) still descend while others
 *                  have leaf tests or new queries to do; the paused descents
 *                  resume in the next step (0 = never pause; default 24).
 *   split_depth    pooled all-hits / crossings / overlap / cast-all calls:
 *                  each query is first walked this many pair-node levels down
 *                  and split into one sub-walk per node it reaches there (or
 *                  leaf it reaches above), so that a few long queries spread
 *                  over the whole chip; 0 = automatic (overlap calls of fewer
 *                  than 2M queries split three quarters of the way down the
 *                  tree, others do not), -1 = never split, 1..24 = fixed.
 *                  Results do not depend on it.
 *   wide           pooled nearest / any-hit calls without a filter: 1 walks
 *                  4-wide records (each pair node with its children's pair
 *                  nodes, rebuilt from the tree before the walk), half the
 *                  dependent node fetches per ray; 0 = the pair nodes; 2 (the
 *                  default) = the 4-wide records for the walks in the
 *                  reference's order (nearest rays over SPHERE / TRIANGLE
 *                  leaves in SQ_HITS_REFERENCE; 0.92x the time on the 8.4M
 *                  heightfield rays, 0.86x on the incoherent half), the pair
 *                  nodes elsewhere.  Over SPHERE / TRIANGLE leaves in
 *                  SQ_HITS_REFERENCE the 4-wide walk needs a tree without
 *                  over-full leaves (coincident centroids); other trees take
 *                  the pair nodes.
 *   lookahead      pooled nearest-ray calls walked in the reference's order
 *                  (SPHERE / TRIANGLE leaves in SQ_HITS_REFERENCE): 1 (the
 *                  default) = a lane that reaches a leaf keeps it pending and
 *                  descends on to its next leaf while the rest of the wave
 *                  descends, then tests it and checks its position again
 *                  against the new limit (the same primitives tested with the
 *                  same limits, results as 0 bit for bit; 0.94x the time on the
 *                  8.4M heightfield rays); 0 = one leaf at a time.
 *   lookahead_pause descend_min of the look-ahead walk (0..64, default 40):
 *                  its lanes keep descending past a pending leaf, so it pauses
 *                  at a higher count than the other walks (on the 8.4M
 *                  heightfield rays 0.96x the time at 40 against 24, where
 *                  the other walks are fastest).  Results do not depend on it.
 *   packet         nearest-hit ray calls without a filter: 1 walks each 64
 *                  rays of the batch (in the Morton order above, for calls of
 *                  >= 4096 rays) as one wave-coherent packet: one walk per
 *                  wave, nodes and primitives loaded once for all lanes, each
 *                  lane testing its own ray against its own limit, so each ray
 *                  visits exactly the nodes of its own walk (results as
 *                  packet 0, bit for bit).  For coherent batches (camera rays:
 *                  1.6x on 4M camera rays of the 2M-triangle heightfield);
 *                  incoherent batches walk many times slower this way.  Not for
 *                  graph capture (the sort). */
typedef struct sq_tuning {
  int block;
  int rays_per_lane;
  int lds_stack;
  int ray_order;
  int descend_min;
  int split_depth;
  int wide;
  int packet;
  int lookahead;
  int lookahead_pause;
} sq_tuning;
int sq_ctx_set_tuning(sq_ctx* ctx, const sq_tuning* tuning);
int sq_ctx_get_tuning(const sq_ctx* ctx, sq_tuning* tuning);
/* Block size only (0 = default). */
int sq_ctx_set_block(sq_ctx* ctx, int block_threads);

/* Time (ms) of the last query kernel on this context, from HIP events recorded
 * on the context stream around it (synchronises). */
int sq_ctx_last_kernel_ms(sq_ctx* ctx, float* ms);

/* ------------------------------------------------------------------------
 * Batches: many nearest-hit batches over many scenes in one persistent,
 * work-stealing launch.  The reference answers one batch per call
 * (raycast_nearest_batch, bvh.rs:134-193, or Tlas::raycast_nearest per ray,
 * tlas.rs:82-154); a batch set turns all batches of a frame into one table of
 * 64 * rays_per_lane-ray chunks that a grid sized to the chip pulls from an
 * atomic counter.  Results equal the single-scene calls (default filter, the
 * given max_toi).  Every scene of a set has the same class, dimension and kind.
 * sq_batch_submit only enqueues kernels (the counter is cleared by a kernel):
 * it may be recorded into a hipGraph on a capturing context stream.  Tuning
 * (block, rays_per_lane, lds_stack) is read at creation.
 * ---------------------------------------------------------------------- */
#define SQ_BATCH_BVH 0    /* scenes are sq_bvh*; out_leaf = primitive id */
#define SQ_BATCH_TLAS 1   /* scenes are sq_tlas*; out_leaf = instance id */
typedef struct sq_batch sq_batch;
typedef struct sq_batch_query {
  const void* scene;      /* sq_bvh* or sq_tlas*, alive while the set is used */
  const float* rays;      /* device float[n][2*dim] */
  uint64_t n;
  float max_toi;
  int32_t* out_leaf;      /* device int32[n], -1 = miss */
  float* out_toi;         /* device float[n] */
} sq_batch_query;
int sq_batch_create(sq_ctx* ctx, int scene_class, const sq_batch_query* queries, uint32_t n_queries, sq_batch** out);
int sq_batch_submit(sq_batch* batch);
/* SQ_OK while every scene of the set keeps the arrays and the stack depth the
 * set was made for (sq_batch_submit checks the same first; a captured graph of
 * submits does not), SQ_ERR_INVALID once one was rebuilt into new arrays or
 * deeper: make the set (and its graph) again. */
int sq_batch_check(sq_batch* batch);
/* Chunks in the set, and the persistent grid (blocks x threads). */
int sq_batch_get_info(const sq_batch* batch, uint32_t* n_chunks, uint32_t* grid, uint32_t* block);
int sq_batch_destroy(sq_batch* batch);

/* ------------------------------------------------------------------------
 * Learned subtree proposal (no counterpart in the reference, which has only the
 * plain walk).  A small MLP maps a ray to the subtrees of the BVH cut at a
 * fixed depth that most likely hold its nearest hit; the proposed subtrees are
 * walked first to tighten t_max and the full exact walk then confirms, so
 * sq_raycast_nearest_proposed returns exactly what sq_raycast_nearest returns
 * under SQ_HITS_GUARDED (whose answer does not depend on the walk order)
 * whatever the network predicts.
 *
 * Network: x = (origin mapped from the scene box to [-1,1]^3, direction);
 * 64 random Fourier features sin/cos(2 pi x.B), B ~ N(0, rff_scale^2) [6][64];
 * 3 x 256 ReLU layers; logits over the subtrees.  BF16 weights (FP32 masters
 * for Adam), FP32 accumulation on the matrix cores.
 * ---------------------------------------------------------------------- */
typedef struct sq_proposal sq_proposal;
/* TRIANGLE scene, cut depth 1..12 (<= 4096 subtrees: the nodes at that depth
 * and the leaves above it), Fourier scale rff_scale > 0, weights and Fourier
 * matrix drawn from seed.  The scene must outlive the proposal. */
int sq_proposal_create(sq_ctx* ctx, sq_bvh* scene, int cut_depth, float rff_scale, uint64_t seed, sq_proposal** out);
/* The same over a two-level scene (D = 3, any instance kind): the subtrees are
 * those of the Tlas's top tree over the instances, cut at cut_depth, the
 * labels the subtrees of the hit instances (sq_proposal_labels takes the
 * instance ids of sq_tlas_raycast_nearest), the scene box the top tree's root
 * box at creation.  The Tlas must outlive the proposal. */
int sq_proposal_create_tlas(sq_ctx* ctx, sq_tlas* scene, int cut_depth, float rff_scale, uint64_t seed, sq_proposal** out);
int sq_proposal_destroy(sq_proposal* p);
int sq_proposal_num_subtrees(const sq_proposal* p, uint32_t* out);
int sq_proposal_num_params(const sq_proposal* p, uint64_t* out);
/* Subtree of each nearest-hit leaf (device int32[n]; -1 for a miss): labels. */
int sq_proposal_labels(sq_proposal* p, const int32_t* leaf, uint64_t n, int32_t* label);
/* One Adam step (lr, 0.9, 0.999, 1e-8) on a batch: rays float[n][6], labels
 * int32[n] (device; label < 0 ignored), n <= 2^24.  loss_out (host, may be
 * NULL): mean softmax cross-entropy over labelled rays (synchronises).
 * grads_out (host, may be NULL): stop after the backward pass, no update, and
 * copy the gradients [dW1, db1, ..., dW4, db4], fp32 [out][in] and [out]. */
int sq_proposal_train_step(sq_proposal* p, const float* rays, const int32_t* labels, uint64_t n, float lr, float* loss_out,
                           float* grads_out);
/* FP32 master weights [W1 [256][128], b1, W2 [256][256], b2, W3, b3,
 * W4 [subtrees][256], b4] (host): copy out (get) and/or replace (set). */
int sq_proposal_weights(sq_proposal* p, float* get, const float* set);
/* The Fourier matrix B float[6][64] and the scene box float[6] (host). */
int sq_proposal_features_spec(const sq_proposal* p, float* rff, float* box);
/* Top-4 subtrees per ray, best first (device uint16[n][4]; 0xFFFF = none). */
int sq_proposal_infer(sq_proposal* p, const float* rays, uint64_t n, uint16_t* top);
/* Exact nearest hit, walking the proposed subtrees (top from
 * sq_proposal_infer; NULL = none, the plain walk) first.  Same results as
 * sq_raycast_nearest under SQ_HITS_GUARDED without per-ray limits or filter
 * (whatever the context's hit mode).  stats (device
 * uint64[3], may be NULL; NULL runs the uncounted kernel) accumulates: rays
 * whose proposed subtrees held a hit, pair nodes fetched, primitives tested. */
int sq_raycast_nearest_proposed(sq_proposal* p, const float* rays, uint64_t n, float max_toi, const uint16_t* top,
                                int32_t* out_leaf, float* out_toi, uint64_t* stats);
/* Exact nearest hit over the Tlas of a proposal from sq_proposal_create_tlas,
 * walking the proposed top-tree subtrees first: the same instance, toi and
 * sub-object (out_sub, may be NULL) as sq_tlas_raycast_nearest under
 * SQ_HITS_GUARDED without per-ray limits (whatever the context's hit mode).
 * stats (device uint64[3], may be NULL): stats[0] accumulates the rays whose
 * proposed subtrees held a hit; stats[1], stats[2] are left unchanged. */
int sq_tlas_raycast_nearest_proposed(sq_proposal* p, const float* rays, uint64_t n, float max_toi, const uint16_t* top,
                                     int32_t* out_inst, float* out_toi, int32_t* out_sub, uint64_t* stats);
/* C[M][N] = bf16(A)[M][K] * bf16(B)[N][K]^T in FP32 on the matrix cores, the
 * GEMM the training step runs (device fp32 inputs, rounded to bf16 first).
 * M, N multiples of 128, K of 32; splits > 1 uses the split-K reduction. */
int sq_debug_gemm(sq_ctx* ctx, const float* A, const float* B, int M, int N, int K, int splits, float* C);

/* ------------------------------------------------------------------------
 * The `f64` feature (crates/spatial-query/src/maths/point.rs:14-18: Scalar is
 * f64): a single-level tree over double-precision SQ_PRIM_AABB or
 * SQ_PRIM_SPHERE primitives, D = 2..6, built with the reference median split
 * in double (tree.rs:112-163) and walked with the same slab and leaf tests as
 * the FP32 tree, in double.  Entries mirror their FP32 counterparts above:
 * same argument meaning, filter, miss values and CSR form; every float
 * argument is double here.  TRIANGLE primitives are refused (the reference's
 * triangle test runs on glam's f32 Vec3, which the feature does not switch).
 * ---------------------------------------------------------------------- */
typedef struct sq_bvh64 sq_bvh64;
/* Bvh::build (bvh.rs:31-37, tree.rs:39-49,112-163) over host double primitives: AABB [min D][max D],
 * SPHERE [centre D][r]. */
int sq_bvh64_create(sq_ctx* ctx, int dim, int kind, const double* prims, uint64_t n, sq_bvh64** out);
int sq_bvh64_destroy(sq_bvh64* bvh);
/* Bvh::refit (bvh.rs:45-48, tree.rs:54-75) under the feature: the same number
 * of host double primitives at new positions (AABB or SPHERE trees); the
 * topology is kept, bounds refit in double, records uploaded again (the
 * device arrays stay where they are).  Synchronises the context stream. */
int sq_bvh64_refit(sq_bvh64* bvh, const double* prims);
/* TreeStats (stats.rs:15-32); bytes counts the f64 node size, quality is the
 * interior surface sum computed in double. */
int sq_bvh64_stats(const sq_bvh64* bvh, sq_tree_stats* out);
/* QueryFilter layers by original leaf id (host or device uint64[n]; NULL clears). */
int sq_bvh64_set_layers(sq_bvh64* bvh, const uint64_t* layers);
/* Bvh::raycast_nearest (bvh.rs:51-124): rays double[n][2D] (device);
 * max_toi_per_ray (device double[n]) may be NULL; out_point / out_normal
 * double[n][D] may be NULL. */
int sq_bvh64_raycast_nearest(sq_bvh64* bvh, const double* rays, uint64_t n, double max_toi,
                             const double* max_toi_per_ray, const sq_filter* filter, int32_t* out_leaf,
                             double* out_toi, double* out_point, double* out_normal);
/* raycast_nearest(..).is_some() per ray. */
int sq_bvh64_raycast_any(sq_bvh64* bvh, const double* rays, uint64_t n, double max_toi, const double* max_toi_per_ray,
                         const sq_filter* filter, uint8_t* out_hit);
/* Bvh::raycast_all (bvh.rs:196-259), CSR as sq_raycast_all. */
int sq_bvh64_raycast_all(sq_bvh64* bvh, const double* rays, uint64_t n, double max_toi, const sq_filter* filter,
                         uint64_t* offsets, int32_t* out_leaf, double* out_toi, uint64_t capacity, uint64_t* total);
/* Bvh::shapecast_nearest / shapecast_all (bvh.rs:331-414): casts double[n][4D]
 * (origin, dir, probe box min, max) as sq_shapecast_nearest / _all;
 * max_toi_per_cast may be NULL.  AABB leaves answer the grown-box slab entry,
 * SPHERE leaves the ball of radius r + probe_max[0]. */
int sq_bvh64_shapecast_nearest(sq_bvh64* bvh, const double* casts, uint64_t n, double max_toi,
                               const double* max_toi_per_cast, const sq_filter* filter, int32_t* out_leaf,
                               double* out_toi, double* out_point);
int sq_bvh64_shapecast_all(sq_bvh64* bvh, const double* casts, uint64_t n, double max_toi,
                           const double* max_toi_per_cast, const sq_filter* filter, uint64_t* offsets,
                           int32_t* out_leaf, double* out_toi, uint64_t capacity, uint64_t* total);
/* Bvh::overlap (bvh.rs:418-456): regions double[n][2D], CSR as sq_overlap. */
int sq_bvh64_overlap(sq_bvh64* bvh, const double* regions, uint64_t n, const sq_filter* filter, uint64_t* offsets,
                     int32_t* out_leaf, uint64_t capacity, uint64_t* total);
/* Tlas over box instances under the `f64` feature (tlas.rs:35-56, the
 * two_level.rs box instances): host rot double[n][D][D] (row-major rotation),
 * trans double[n][D], halves double[n][D] (local box [-half, half]); D = 2..6.
 * World boxes are Isometry::transform_aabb (isometry.rs:84-97) in double. */
typedef struct sq_tlas64 sq_tlas64;
int sq_tlas64_create(sq_ctx* ctx, int dim, uint64_t n, const double* rot, const double* trans, const double* halves,
                     sq_tlas64** out);
/* Tlas::build over f64 instances of a built-in kind: SQ_INST_BOX (halves
 * double[n][D], as sq_tlas64_create) or SQ_INST_BALL (D = 3, radii double[n]:
 * the reference's BallInstances, tests/shape_overlap.rs:288-345).  Host
 * pointers. */
int sq_tlas64_create_kind(sq_ctx* ctx, int dim, int inst_kind, uint64_t n, const double* rot, const double* trans,
                          const double* halves, sq_tlas64** out);
int sq_tlas64_destroy(sq_tlas64* tlas);
/* Tlas::refit (tlas.rs:49-54) under the feature: new host transforms, and
 * half extents (BOX) or radii (BALL), or NULL to keep them; world boxes by
 * transform_aabb in double, the top tree refit.  Synchronises. */
int sq_tlas64_refit(sq_tlas64* tlas, const double* rot, const double* trans, const double* halves);
/* Tlas::needs_rebuild (tlas.rs:58-60): interior surface sum now > factor
 * times its value at the build. */
int sq_tlas64_needs_rebuild(sq_tlas64* tlas, double factor, int* out);
int sq_tlas64_stats(const sq_tlas64* tlas, sq_tree_stats* out);
/* Tlas::raycast_nearest (tlas.rs:104-154) as sq_tlas_raycast_nearest: the
 * instance id and toi of the nearest hit (out_point double[n][D] may be NULL). */
int sq_tlas64_raycast_nearest(sq_tlas64* tlas, const double* rays, uint64_t n, double max_toi,
                              const double* max_toi_per_ray, int32_t* out_inst, double* out_toi, double* out_point);
int sq_tlas64_raycast_any(sq_tlas64* tlas, const double* rays, uint64_t n, double max_toi,
                          const double* max_toi_per_ray, uint8_t* out_hit);
/* Tlas::raycast_all (tlas.rs:158-215) in double, CSR as sq_bvh64_raycast_all
 * (instances by (toi, instance)); also Tlas::raycast_crossings (tlas.rs:221-304)
 * for these kinds, which report one crossing per instance. */
int sq_tlas64_raycast_all(sq_tlas64* tlas, const double* rays, uint64_t n, double max_toi, uint64_t* offsets,
                          int32_t* out_inst, double* out_toi, uint64_t capacity, uint64_t* total);
/* Tlas::overlap (tlas.rs:393-432) in double: every instance of a leaf node
 * touching the region that passes test_overlap (exact for balls, true for
 * boxes), instances ascending. */
int sq_tlas64_overlap(sq_tlas64* tlas, const double* regions, uint64_t n, uint64_t* offsets, int32_t* out_inst,
                      uint64_t capacity, uint64_t* total);
/* Tlas::shapecast_nearest / shapecast_all (tlas.rs:306-391) in double: casts
 * double[n][4*D] (origin, dir, probe box min, max).  Ball instances answer the
 * cast ray against the ball of radius r + the probe's max x about the centre;
 * box instances keep the trait default (no contact: every cast misses). */
int sq_tlas64_shapecast_nearest(sq_tlas64* tlas, const double* casts, uint64_t n, double max_toi, int32_t* out_inst,
                                double* out_toi, double* out_point);
int sq_tlas64_shapecast_all(sq_tlas64* tlas, const double* casts, uint64_t n, double max_toi, uint64_t* offsets,
                            int32_t* out_inst, double* out_toi, uint64_t capacity, uint64_t* total);

/* Consumer narrow phases in double: any QueryGeometry provider under the f64
 * feature (geometry.rs:33-43 test_ray with Scalar = f64), as sq_narrow_create
 * above does it for FP32.  The source defines
 *   __device__ bool sq_narrow64(const double* prim, const double* origin,
 *                               const double* dir, double max_toi, double* toi);
 * over prim_doubles doubles per primitive (SQ_RTC_D, SQ_RTC_PF as above) and is
 * compiled at run time (hipRTC) into the pooled f64 walk.  A hit counts in
 * [0, max_toi].  Nearest queries return the reference's answer (its LIFO order
 * with the consumer's unguarded hits); under SQ_HITS_GUARDED a hit is kept only
 * inside the slab interval of its world box and the walk is near-first.  All
 * hits are every hit in both modes (CSR as sq_bvh64_raycast_all). */
typedef struct sq_narrow64 sq_narrow64;
int sq_narrow64_create(sq_ctx* ctx, int dim, const char* source, uint32_t prim_doubles, sq_narrow64** out);
int sq_narrow64_destroy(sq_narrow64* narrow);
/* The f64 tree over n consumer primitives: host world boxes double[n][2*dim]
 * (QueryGeometry::world_aabb, the median split's input) and host
 * double[n][prim_doubles] (what sq_narrow64 sees).  Only the consumer entries
 * below answer it (the sq_bvh64 queries refuse it; stats and destroy as any
 * sq_bvh64). */
int sq_bvh64_build_consumer(sq_narrow64* narrow, const double* boxes, const double* prims, uint64_t n, sq_bvh64** out);
/* Bvh::raycast_nearest_batch (bvh.rs:134-193) with the consumer's test_ray, in
 * double: rays double[n][2*dim], out_leaf -1 and out_toi +inf on a miss. */
int sq_raycast64_nearest_consumer(sq_narrow64* narrow, sq_bvh64* bvh, const double* rays, uint64_t n, double max_toi,
                                  int32_t* out_leaf, double* out_toi);
/* Bvh::raycast_all (bvh.rs:196-259) with the consumer's test_ray, in double. */
int sq_raycast64_all_consumer(sq_narrow64* narrow, sq_bvh64* bvh, const double* rays, uint64_t n, double max_toi,
                              uint64_t* offsets, int32_t* out_leaf, double* out_toi, uint64_t capacity, uint64_t* total);
/* Bvh::raycast_crossings (bvh.rs:264-328) under the feature: every crossing
 * the source reports through an optional
 *   #define SQ_NARROW64_CROSSINGS 1
 *   template <class Out> __device__ void sq_narrow64_crossings(const double* prim,
 *       const double* origin, const double* dir, double max_toi, Out& out);
 * (out(toi) once per crossing, any order; each kept in [0, max_toi]), or
 * without one the sq_narrow64 hit (the trait's default); CSR as
 * sq_raycast64_all_consumer. */
int sq_raycast64_crossings_consumer(sq_narrow64* narrow, sq_bvh64* bvh, const double* rays, uint64_t n, double max_toi,
                                    uint64_t* offsets, int32_t* out_leaf, double* out_toi, uint64_t capacity,
                                    uint64_t* total);

#ifdef __cplusplus
}
#endif
#endif /* SQ_H */
