/*
 * Mesh C-ABI of the MI355X build (libmesh.so -> libstream_amd.so).
 *
 * The first block is the reference's alpha-shape interface, same names,
 * argument meaning and fp_tt = float signatures; each line cites the
 * reference declaration it replaces.  The second block adds FP64 nodes and
 * the reference's C++-only queries.  Status-returning extensions return 0
 * on success; the message of a failure is in Stream_last_error().  The
 * reference's void entry points record their status on the handle
 * (AlphaShape2D_last_status).
 */
#ifndef STREAM_AMD_MESH_H
#define STREAM_AMD_MESH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct AlphaShape2D AlphaShape2D;

/* include/mesh/DirectAlphaShape2D.hpp:48 */
AlphaShape2D* AlphaShape2D_create(void);
/* include/mesh/DirectAlphaShape2D.hpp:49 */
void AlphaShape2D_destroy(AlphaShape2D* ashape);
/* include/mesh/DirectAlphaShape2D.hpp:53  coords row-major [2N], alpha [N] */
void AlphaShape2D_set_nodes(AlphaShape2D* ashape, uint32_t nnodes, const float* coords, const float* alpha);
/* include/mesh/DirectAlphaShape2D.hpp:56  LBVH + infinity nodes + allocation */
void AlphaShape2D_init(AlphaShape2D* ashape);
/* include/mesh/DirectAlphaShape2D.hpp:59  compute + compress */
void AlphaShape2D_compute(AlphaShape2D* ashape);
/* include/mesh/DirectAlphaShape2D.hpp:62 */
uint32_t AlphaShape2D_get_nelem(const AlphaShape2D* ashape);
/* include/mesh/DirectAlphaShape2D.hpp:65  elems [3*nelem], Morton ids */
void AlphaShape2D_get_elem(const AlphaShape2D* ashape, uint32_t* elems);
/* include/mesh/DirectAlphaShape2D.hpp:67  nodes [2N], Morton order */
void AlphaShape2D_get_ordered_nodes(AlphaShape2D* ashape, float* nodes);

/* ---- extensions ---- */
/* Morton bits per axis (1..32, default 32 as lbvh.hpp:186) for the next init. */
int32_t AlphaShape2D_set_morton_bits(AlphaShape2D* ashape, int32_t bits);
/* FP64 nodes and exact FP64 predicates: coords [2N], alpha [N]. */
int32_t AlphaShape2D_set_nodes_f64(AlphaShape2D* ashape, uint32_t nnodes, const double* coords, const double* alpha);
/* FP64 nodes already in device memory: rows (x, y, unused, alpha). */
int32_t AlphaShape2D_set_nodes_device_f64(AlphaShape2D* ashape, uint32_t nnodes, const double* d_xypr);
int32_t AlphaShape2D_get_ordered_nodes_f64(AlphaShape2D* ashape, double* nodes);
/* number of directed edges; AlphaShape2D::getEdge, src/mesh/DirectAlphaShape2D.cpp:90-98 */
uint32_t AlphaShape2D_get_nedge(const AlphaShape2D* ashape);
/* edge CSR row [N+1], neig [nedge], edge_is_bnd [nedge], node_is_bnd [N];
 * getEdge / getBoundaryNodes, DirectAlphaShape2D.cpp:90-105.  Any may be NULL. */
int32_t AlphaShape2D_get_edges(const AlphaShape2D* ashape, uint32_t* row, uint32_t* neig, uint8_t* edge_is_bnd,
                               uint8_t* node_is_bnd);
/* perm[i] = input index of Morton slot i; getPermutation, DirectAlphaShape2D.cpp:75-80 */
int32_t AlphaShape2D_get_permutation(const AlphaShape2D* ashape, uint32_t* perm);
/* Nodal fields between input order and Morton order: permuteForward /
 * permuteBackward (DirectAlphaShape2D.hpp:220-242), forward out[i] =
 * in[perm[i]], backward out[perm[i]] = in[i]; nnodes elements of
 * elem_bytes each.  Host arrays, or device buffers (_device: on the
 * library stream, not synchronised; in and out must not overlap). */
int32_t AlphaShape2D_permute_forward(const AlphaShape2D* ashape, const void* in, void* out, uint32_t elem_bytes);
int32_t AlphaShape2D_permute_backward(const AlphaShape2D* ashape, const void* in, void* out, uint32_t elem_bytes);
int32_t AlphaShape2D_permute_forward_device(const AlphaShape2D* ashape, const void* d_in, void* d_out,
                                            uint32_t elem_bytes);
int32_t AlphaShape2D_permute_backward_device(const AlphaShape2D* ashape, const void* d_in, void* d_out,
                                             uint32_t elem_bytes);
/* nodes whose local triangulation needed the exact predicate stage */
uint32_t AlphaShape2D_get_nexact(const AlphaShape2D* ashape);
/* status of the last call on the handle (0 = ok) */
int32_t AlphaShape2D_last_status(const AlphaShape2D* ashape);
/* no reference counterpart (the timed loop of testcases/IMR26/uniform2D.cpp:
 * 159-183): init + compute with ms [8] = LBVH build, candidate search +
 * local Delaunay + filter + scans, compaction (GPU times), host wall time of
 * the two calls, GPU time from build start to compaction end; then, within
 * the second, the candidate search, the local Delaunay of every node and the
 * exact redo of undecided nodes */
int32_t AlphaShape2D_timed_run(AlphaShape2D* ashape, double* ms);

/* ======================== 3D alpha-shape ======================== */
typedef struct AlphaShape3D AlphaShape3D;

/* include/mesh/DirectAlphaShape3D.hpp:47 */
AlphaShape3D* AlphaShape3D_create(void);
/* include/mesh/DirectAlphaShape3D.hpp:48 */
void AlphaShape3D_destroy(AlphaShape3D* ashape);
/* include/mesh/DirectAlphaShape3D.hpp:52, src/mesh/DirectAlphaShape3D.cpp:760-772
 * coords row-major [3N], alpha [N] */
void AlphaShape3D_set_nodes(AlphaShape3D* ashape, uint32_t nnodes, const float* coords, const float* alpha);
/* include/mesh/DirectAlphaShape3D.hpp:55 */
void AlphaShape3D_init(AlphaShape3D* ashape);
/* include/mesh/DirectAlphaShape3D.hpp:58  compute + compress */
void AlphaShape3D_compute(AlphaShape3D* ashape);
/* include/mesh/DirectAlphaShape3D.hpp:61 */
uint32_t AlphaShape3D_get_nelem(const AlphaShape3D* ashape);
/* include/mesh/DirectAlphaShape3D.hpp:64  tetrahedra [4*nelem], Morton ids */
void AlphaShape3D_get_elem(const AlphaShape3D* ashape, uint32_t* elems);
/* include/mesh/DirectAlphaShape3D.hpp:66  nodes [3N], Morton order */
void AlphaShape3D_get_ordered_nodes(AlphaShape3D* ashape, float* nodes);

/* ---- extensions ---- */
/* Morton bits per axis (1..21, default 21 as lbvh.hpp:186 for dim 3). */
int32_t AlphaShape3D_set_morton_bits(AlphaShape3D* ashape, int32_t bits);
int32_t AlphaShape3D_set_nodes_f64(AlphaShape3D* ashape, uint32_t nnodes, const double* coords, const double* alpha);
/* FP64 nodes already in device memory: rows (x, y, z, alpha). */
int32_t AlphaShape3D_set_nodes_device_f64(AlphaShape3D* ashape, uint32_t nnodes, const double* d_xyzr);
int32_t AlphaShape3D_get_ordered_nodes_f64(AlphaShape3D* ashape, double* nodes);
/* AlphaShape3D::getEdge / getBoundaryNodes / getPermutation,
 * src/mesh/DirectAlphaShape3D.cpp:75-105; edge_is_bnd = count <= 3 */
uint32_t AlphaShape3D_get_nedge(const AlphaShape3D* ashape);
int32_t AlphaShape3D_get_edges(const AlphaShape3D* ashape, uint32_t* row, uint32_t* neig, uint8_t* edge_is_bnd,
                               uint8_t* node_is_bnd);
int32_t AlphaShape3D_get_permutation(const AlphaShape3D* ashape, uint32_t* perm);
/* permuteForward / permuteBackward of DirectAlphaShape3D.hpp (as the 2D ones) */
int32_t AlphaShape3D_permute_forward(const AlphaShape3D* ashape, const void* in, void* out, uint32_t elem_bytes);
int32_t AlphaShape3D_permute_backward(const AlphaShape3D* ashape, const void* in, void* out, uint32_t elem_bytes);
int32_t AlphaShape3D_permute_forward_device(const AlphaShape3D* ashape, const void* d_in, void* d_out,
                                            uint32_t elem_bytes);
int32_t AlphaShape3D_permute_backward_device(const AlphaShape3D* ashape, const void* d_in, void* d_out,
                                             uint32_t elem_bytes);
/* nodes that needed the exact stage; nodes whose candidate list was clamped
 * to 124 (the reference's 128 entries with the 4 infinity nodes, cpp:369) */
uint32_t AlphaShape3D_get_nexact(const AlphaShape3D* ashape);
uint32_t AlphaShape3D_get_nclamped(const AlphaShape3D* ashape);
int32_t AlphaShape3D_last_status(const AlphaShape3D* ashape);
/* the same for testcases/IMR26/uniform3D.cpp */
int32_t AlphaShape3D_timed_run(AlphaShape3D* ashape, double* ms);

/* Full 2D Delaunay (include/mesh/mesh.hpp:23-62, src/mesh/mesh.cpp:800-862),
 * float nodes.  Triangles and node ids are in Morton order, as in the
 * reference; remove_super_nodes must precede get_nelem for a non-empty
 * result. */
typedef struct Mesh2D Mesh2D;

/* include/mesh/mesh.hpp:39 */
Mesh2D* Mesh2D_create(void);
/* include/mesh/mesh.hpp:42 */
void Mesh2D_destroy(Mesh2D* mesh);
/* include/mesh/mesh.hpp:45  nodes row-major [2N] (N >= 2) */
void Mesh2D_set_nodes(Mesh2D* mesh, uint32_t nnodes, const float* coords);
/* include/mesh/mesh.hpp:48  agglomerative BVH, super-triangle, 3-triangle stars */
void Mesh2D_init(Mesh2D* mesh);
/* include/mesh/mesh.hpp:51  Morton neighbours (+-5) */
void Mesh2D_insert_morton_neighbors(Mesh2D* mesh);
/* include/mesh/mesh.hpp:55  neighbours whose BVH kd cell touches the node's */
void Mesh2D_insert_BVH_neighbors(Mesh2D* mesh);
/* include/mesh/mesh.hpp:58  up to 32 rounds of circumcircle search + insertion */
void Mesh2D_insert_iterative(Mesh2D* mesh);
/* include/mesh/mesh.hpp:60 */
void Mesh2D_remove_super_nodes(Mesh2D* mesh);
/* include/mesh/mesh.hpp:62  side effect: compresses the stars (mesh.cpp:849-852) */
uint32_t Mesh2D_get_nelem(Mesh2D* mesh);
/* include/mesh/mesh.hpp:63  [nelem][3] */
void Mesh2D_get_elem(Mesh2D* mesh, uint32_t* elems);
/* src/mesh/mesh.cpp:860 (defined and bound there, not declared in mesh.hpp)  [N][2] */
void Mesh2D_get_ordered_nodes(Mesh2D* mesh, float* nodes);
/* status of the last void call on the handle */
int32_t Mesh2D_last_status(const Mesh2D* mesh);
/* counters [3]: nodes frozen on a capacity overflow (64 triangles / local
 * indices / heap entries per node), most rounds of insert_iterative, exact
 * incircle evaluations; ms [5]: GPU time of init, Morton, BVH, iterative,
 * remove phases.  Either may be NULL. */
int32_t Mesh2D_get_stats(Mesh2D* mesh, uint32_t* counters, float* ms);
/* stars: elem [N][64] (a | b << 8), nelem [N], neig [N][64] (local index
 * j >= 3 at j - 3), nlocal [N] (bit 31 = frozen), Morton permutation [N] */
int32_t Mesh2D_get_stars(Mesh2D* mesh, uint16_t* elem, uint32_t* nelem, uint32_t* neig, uint32_t* nlocal,
                         uint32_t* perm);
/* kd cells of the 2N-1 BVH nodes, rows (lo0, lo1, hi0, hi1), after
 * insert_BVH_neighbors */
int32_t Mesh2D_get_cells(Mesh2D* mesh, float* cells);

#ifdef __cplusplus
}
#endif
#endif
