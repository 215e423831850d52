/*
 * Geometry and runtime C-ABI of the MI355X build (libgeometry.so ->
 * libstream_amd.so).
 *
 * The reference's libgeometry exports no C symbol (src/geometry/lbvh.cpp:
 * 22-27 only instantiates templates); these entry points expose the GPU
 * LBVH of include/geometry/lbvh.hpp:61-362 (Morton codes, stable sort,
 * Karras topology, refit) and the device primitives that replace AVA's
 * sort/scan (include/geometry/lbvh.hpp:217-235,
 * src/mesh/DirectAlphaShape2D.cpp:228-281) for tests and tools.  Host
 * arrays in and out; 0 = success.
 */
#ifndef STREAM_AMD_GEOMETRY_H
#define STREAM_AMD_GEOMETRY_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* Stream_last_error(void);
int32_t Stream_device_count(void);
int32_t Stream_synchronize(void);
/* the reference benchmarks' point generator (testcases/IMR26/uniform2D.cpp:
 * 130-141, uniform3D.cpp): srand(seed), (float)rand()/RAND_MAX per
 * coordinate; out [n][dim], dim 2 or 3 (host only) */
int32_t Stream_uniform_cloud_f32(uint32_t n, uint32_t dim, uint32_t seed, float* out);
/* no reference counterpart: device bytes held by the library now / at the
 * high-water mark (reset_peak != 0 restarts it), device free / total memory */
int32_t Stream_device_bytes(uint64_t* current, uint64_t* peak, int32_t reset_peak, uint64_t* dev_free,
                            uint64_t* dev_total);

/* xyr: rows (x, y, r).  Outputs may be NULL.  ibox rows (xmin, ymin, xmax,
 * ymax) per internal node; gbox the global sphere box.  leaf_mode 0 boxes
 * centres (alpha-shape functor, DirectAlphaShape2D.hpp:87-99), 1 boxes
 * spheres (ComputeBB_Functor, primitives.hpp:307-322). */
int32_t LBVH2D_build_f32(uint32_t n, const float* xyr, int32_t bits, int32_t leaf_mode, uint64_t* codes_sorted,
                         uint32_t* perm, int32_t* sep, uint8_t* child_is_leaf, uint32_t* leaf_parent,
                         uint32_t* internal_parent, float* ibox, float* gbox);
int32_t LBVH2D_build_f64(uint32_t n, const double* xyr, int32_t bits, int32_t leaf_mode, uint64_t* codes_sorted,
                         uint32_t* perm, int32_t* sep, uint8_t* child_is_leaf, uint32_t* leaf_parent,
                         uint32_t* internal_parent, double* ibox, double* gbox);
int32_t LBVH3D_build_f64(uint32_t n, const double* xyzr, int32_t bits, int32_t leaf_mode, uint64_t* codes_sorted,
                         uint32_t* perm, int32_t* sep, uint8_t* child_is_leaf, uint32_t* leaf_parent,
                         uint32_t* internal_parent, double* ibox, double* gbox);
/* unsorted Morton codes, lbvh.hpp:178-213 */
int32_t LBVH2D_morton_f32(uint32_t n, const float* xyr, int32_t bits, uint64_t* codes);
int32_t LBVH2D_morton_f64(uint32_t n, const double* xyr, int32_t bits, uint64_t* codes);
int32_t LBVH3D_morton_f64(uint32_t n, const double* xyzr, int32_t bits, uint64_t* codes);
/* sphere-sphere overlaps (self excluded), tests/geometry/lbvh_simple_collision.cpp:100-140 */
int32_t LBVH2D_collisions_f32(uint32_t n, const float* xyr, int32_t bits, uint64_t* total);
int32_t LBVH2D_collisions_f64(uint32_t n, const double* xyr, int32_t bits, uint64_t* total);

/* Agglomerative LBVH (include/geometry/lbvh_agglo.hpp:62-361, LBVHa::build):
 * permutation [n], children [n-1][2], ranges [n-1][2] (internal node k has
 * id k + n), parent [2n-1] (2n at the root), root, node boxes [2n-1] rows
 * (min..., max...), split index [n-1].  Any output may be NULL. */
int32_t LBVHa2D_build_f32(uint32_t n, const float* xyr, int32_t bits, int32_t leaf_mode, uint32_t* perm,
                          uint32_t* children, uint32_t* range, uint32_t* parent, uint32_t* root, float* boxes,
                          uint32_t* split);
int32_t LBVHa2D_build_f64(uint32_t n, const double* xyr, int32_t bits, int32_t leaf_mode, uint32_t* perm,
                          uint32_t* children, uint32_t* range, uint32_t* parent, uint32_t* root, double* boxes,
                          uint32_t* split);
int32_t LBVHa3D_build_f64(uint32_t n, const double* xyzr, int32_t bits, int32_t leaf_mode, uint32_t* perm,
                          uint32_t* children, uint32_t* range, uint32_t* parent, uint32_t* root, double* boxes,
                          uint32_t* split);
/* overlaps with leaf buckets of up to leaf_size objects,
 * tests/geometry/lbvh_agglo_simple_collision.cpp:100-165 */
int32_t LBVHa2D_collisions_f32(uint32_t n, const float* xyr, int32_t bits, uint32_t leaf_size, uint64_t* total);
int32_t LBVHa2D_collisions_f64(uint32_t n, const double* xyr, int32_t bits, uint32_t leaf_size, uint64_t* total);
int32_t LBVHa3D_collisions_f64(uint32_t n, const double* xyzr, int32_t bits, uint32_t leaf_size, uint64_t* total);

/* TRBVH in float (include/geometry/trbvh.hpp:64-681, TRBVH::build,
 * :637-679): the agglomerative topology with surface-area costs, then
 * `rounds` treelet restructuring passes (min subtree 5, 10, 20, ...; the
 * reference runs 3).  Outputs as LBVHa2D_build_* (range as built; the
 * restructuring leaves it stale, as the reference does) plus node costs
 * sah [2n-1] and subtree object counts [2n-1]; ms (NULL or [2]) gets the
 * GPU time of the build and of the rounds.  Any output may be NULL. */
int32_t TRBVH2D_build_f32(uint32_t n, const float* xyr, int32_t rounds, uint32_t* perm, uint32_t* children,
                          uint32_t* range, uint32_t* parent, uint32_t* root, float* boxes, float* sah,
                          uint32_t* subtree, double* ms);
int32_t TRBVH3D_build_f32(uint32_t n, const float* xyzr, int32_t rounds, uint32_t* perm, uint32_t* children,
                          uint32_t* range, uint32_t* parent, uint32_t* root, float* boxes, float* sah,
                          uint32_t* subtree, double* ms);
/* overlaps (self excluded), tests/geometry/trbvh_random_collision.cpp:87-139 */
int32_t TRBVH2D_collisions_f32(uint32_t n, const float* xyr, int32_t rounds, uint64_t* total);
int32_t TRBVH3D_collisions_f32(uint32_t n, const float* xyzr, int32_t rounds, uint64_t* total);

/* device TraversalMinHeap (include/geometry/traversal_stack.hpp:31-90): push
 * n items (32 slots), pop all into out_idx/out_cost; *pushed = accepted. */
int32_t Stream_heap_order(uint32_t n, const uint32_t* idx, const float* cost, uint32_t* out_idx, float* out_cost,
                          int32_t* pushed);

/* no reference counterpart (verification hook): the exact sign of orient2d
 * (kind 0, 3 points x 2), incircle (1, 4 x 2), orient3d (2, 4 x 3) or
 * insphere (3, 5 x 3) by the device exact stage compiled for the host; pts
 * float (fp64 = 0) or double (predicates.hpp:342-1100 semantics) */
int32_t Stream_exact_sign_host(int32_t kind, int32_t fp64, const void* pts, int32_t* sign);

/* stable onesweep radix sort of (u64, u32) pairs over the low `bits` bits */
int32_t Stream_sort_pairs_u64(uint32_t n, const uint64_t* keys, const uint32_t* vals, uint64_t* keys_out,
                              uint32_t* vals_out, int32_t bits);
/* plain copy of nbytes (multiple of 16) repeated reps times; average ms.
 * Calibration kernel for hardware byte counters. */
int32_t Stream_copy_calibration(uint64_t nbytes, uint32_t reps, double* ms);
/* exclusive scans, out [n+1] with out[n] = total */
int32_t Stream_exclusive_scan_u32(uint32_t n, const uint32_t* in, uint32_t* out);
int32_t Stream_exclusive_scan_u8(uint32_t n, const uint8_t* in, uint32_t* out);
/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (serving loops, harnesses that forbid hipMalloc): the scan
 * of Stream_exclusive_scan_u32 on caller-owned DEVICE buffers, in [n] and
 * out [n+1], with its look-back workspace in a static per-device scratch
 * array; n <= 2^28 (larger n is refused, no launch).  Runs on `stream` (a
 * hipStream_t, NULL = the null stream) and waits for it.  Makes no device
 * allocation; calls that overlap on one device share the scratch and race. */
int32_t Stream_exclusive_scan_u32_noalloc(uint32_t n, const uint32_t* d_in, uint32_t* d_out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
