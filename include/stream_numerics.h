/*
 * Numerics C-ABI of the MI355X build (libnumerics.so -> libstream_amd.so).
 *
 * First block: the reference's CSR / linear-system / Jacobi / CG interface
 * with its fp_tt = float signatures and semantics (SolverCG_* stop on an
 * absolute 1e-6 on sqrt(r.z) or sqrt(r.r), at most 10000 iterations, and
 * return the reference's post-incremented count).  Second block: FP64
 * values, the FP64 relative-residual and mixed-precision solvers, and the
 * mesh-and-solve step, which has no reference counterpart.
 */
#ifndef STREAM_AMD_NUMERICS_H
#define STREAM_AMD_NUMERICS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct d_CSR d_CSR;
typedef struct h_CSR h_CSR;
typedef struct LinSys LinSys;
typedef struct PrecJacobi_st PrecJacobi_st;
typedef struct SolverCG SolverCG;
typedef struct PfemStep2D PfemStep2D;

/* include/numerics/csr.hpp:51-68 */
d_CSR* d_csr_create(void);
void d_csr_destroy(d_CSR* A);
h_CSR* h_csr_create(uint32_t n, const uint32_t* row, const uint32_t* col, const float* val);
void h_csr_destroy(h_CSR* A);
void d_csr_h2d(d_CSR* dcsr, const h_CSR* hcsr);
/* include/numerics/linear_system.hpp:48-54 */
LinSys* LinSys_create(void);
void LinSys_set(LinSys* sys, const d_CSR* A, const float* b);
void LinSys_destroy(LinSys* sys);
/* include/numerics/preconditioner.hpp:39-40 */
PrecJacobi_st* PrecJacobi_create(const d_CSR* A);
void PrecJacobi_destroy(PrecJacobi_st* prec);
/* include/numerics/solver.hpp:54-61 */
SolverCG* SolverCG_create(void);
void SolverCG_destroy(SolverCG* solver);
uint32_t SolverCG_jacobi_solve(SolverCG* solver, const LinSys* sys, const PrecJacobi_st* prec, float* x);
uint32_t SolverCG_solve(SolverCG* solver, const LinSys* sys, float* x);

/* ---- extensions ---- */
h_CSR* h_csr_create_f64(uint32_t n, const uint32_t* row, const uint32_t* col, const double* val);
int32_t LinSys_set_f64(LinSys* sys, const d_CSR* A, const double* b);
/* rows without a usable diagonal (applied as identity) */
uint32_t PrecJacobi_missing_diagonals(const PrecJacobi_st* prec);
/* FP64 Jacobi-PCG to ||b-Ax||/||b|| <= rtol; stats [4]: iterations, true
 * relative residual, ms, residual replacements.  Non-zero status when the
 * residual is not reached. */
int32_t SolverCG_solve_rel_f64(SolverCG* solver, const LinSys* sys, const PrecJacobi_st* prec, double* x, double rtol,
                               uint32_t maxit, double* stats);
/* FP32 inner iterations, FP64 dots and FP64 residual replacement; same stats. */
int32_t SolverCG_solve_mixed(SolverCG* solver, const LinSys* sys, const PrecJacobi_st* prec, double* x, double rtol,
                             uint32_t maxit, double* stats);

/* Process-wide solver knobs: "replace_factor" (mixed precision residual
 * replacement threshold, default 0.1), "batch" (iterations between host
 * checks, default 32), "cg_blocks" (grid cap of the CG kernels, 64..4096,
 * default 2048), "graph" (1: replay each batch as a captured HIP graph,
 * default 0), "xcd_map" (SELL slices in XCD-contiguous chunks, default 1)
 * "fold_group" (blocks per group of the two-level reduction, 0 = one
 * level; 8, 16, 32 (default) or 64), "sell_chunk" (SELL entries per load
 * step of the CG row loops, 4 or 8 (default)) and "llt_threads" (block size
 * of the L L^T path's kernels, 256, 512 (default) or 1024; "cg_blocks"
 * counts 256-thread blocks, so the thread count is the same). */
int32_t Stream_tune(const char* key, double value);

/* Mesh-and-solve step on device-resident points: LBVH + 2D alpha-shape +
 * P1 pressure assembly (Dirichlet p = 0 on free-surface nodes) + CG. */
PfemStep2D* PfemStep2D_create(void);
void PfemStep2D_destroy(PfemStep2D* step);
/* xy [2N], alpha [N], nodal source f [N], input order */
int32_t PfemStep2D_set_points(PfemStep2D* step, uint32_t n, const double* xy, const double* alpha, const double* f);
int32_t PfemStep2D_set_morton_bits(PfemStep2D* step, int32_t bits);
/* solver 0 = FP64 Jacobi-PCG, 1 = mixed FP32/FP64 Jacobi-PCG, 2 = mixed PCG
 * with the FSAI L L^T factor, 3 = mixed PCG with the learned L L^T factor
 * (stream_learned.h; Jacobi when the factor is rejected).  stats [10]:
 * mesh ms, assembly ms, solve ms, iterations, true relative residual,
 * replacements, nnz, triangles, preconditioner set-up ms, 1 if the learned
 * factor was rejected. */
int32_t PfemStep2D_run(PfemStep2D* step, int32_t solver, double rtol, uint32_t maxit, double* stats);
/* Mesh + assemble only; stats [2] (nullable): mesh ms, assembly ms. */
int32_t PfemStep2D_assemble(PfemStep2D* step, double* stats);
/* Device source f = sum (ax kx + ay ky) cos(kx x + ky y + phi), modes
 * [n][5] rows (kx, ky, phi, ax, ay), n <= 32, re-evaluated after advection. */
int32_t PfemStep2D_set_source_modes(PfemStep2D* step, uint32_t n_modes, const double* modes);
/* `steps` forward-Euler steps of u = (dpsi/dy, -dpsi/dx), psi = sum a sin(kx x
 * + ky y + phi), modes [n][4] rows (kx, ky, phi, a), n <= 32. */
int32_t PfemStep2D_advect(PfemStep2D* step, uint32_t n_modes, const double* modes, double dt, uint32_t steps);
/* current points [N][2], input order */
int32_t PfemStep2D_get_points(PfemStep2D* step, double* xy);
/* no reference counterpart (the PFEM velocity step named as a goal in the
 * reference README.md:12-18): nodal velocity [N][2] in input order, one
 * fractional (projection) step -- predictor u* = u + dt (g - nu M^-1 K u),
 * pressure K p = -(rho/dt) D(u*) with p = 0 on free-surface nodes solved by
 * `solver` (as PfemStep2D_run) to rtol, correction u = u* - (dt/rho) M^-1
 * G(p), then (move != 0) the points move by dt u; g [2]; stats [11] =
 * PfemStep2D_run's ten + ms of the velocity kernels -- and the last step's
 * fields in Morton order (u*, new velocity, lumped mass, D(u_new)) */
int32_t PfemStep2D_set_velocity(PfemStep2D* step, const double* uv);
int32_t PfemStep2D_get_velocity(PfemStep2D* step, double* uv);
int32_t PfemStep2D_projection_step(PfemStep2D* step, int32_t solver, double dt, double rho, double nu, const double* g,
                                   double rtol, uint32_t maxit, int32_t move, double* stats);
int32_t PfemStep2D_get_flow_fields(PfemStep2D* step, double* us, double* unew, double* mass, double* div);

/* 3D mesh-and-solve step: LBVH + alpha-shape tetrahedra + P1 pressure
 * assembly on tetrahedra + CG (solvers 0, 1, 2 as above). */
typedef struct PfemStep3D PfemStep3D;
PfemStep3D* PfemStep3D_create(void);
void PfemStep3D_destroy(PfemStep3D* step);
/* xyz [3N], alpha [N], nodal source f [N], input order */
int32_t PfemStep3D_set_points(PfemStep3D* step, uint32_t n, const double* xyz, const double* alpha, const double* f);
int32_t PfemStep3D_set_morton_bits(PfemStep3D* step, int32_t bits);
int32_t PfemStep3D_run(PfemStep3D* step, int32_t solver, double rtol, uint32_t maxit, double* stats);
/* no reference counterpart: forward-Euler advection of the points in the
 * divergence-free field u(x) = sum cos(k . x + phi) b of n_modes rows
 * (kx, ky, kz, phi, bx, by, bz), and the current points (input order) */
int32_t PfemStep3D_advect(PfemStep3D* step, uint32_t n_modes, const double* modes, double dt, uint32_t steps);
int32_t PfemStep3D_get_points(PfemStep3D* step, double* xyz);
int32_t PfemStep3D_assemble(PfemStep3D* step, double* stats);
int32_t PfemStep3D_get_solution(PfemStep3D* step, double* p);
uint32_t PfemStep3D_nnz(PfemStep3D* step);
uint32_t PfemStep3D_get_nelem(PfemStep3D* step);
int32_t PfemStep3D_get_system(PfemStep3D* step, uint32_t* row, uint32_t* col, double* val, double* b);
int32_t PfemStep3D_get_mesh(PfemStep3D* step, uint32_t* elem, uint8_t* node_is_bnd, double* xyz, uint32_t* perm);
int32_t PfemStep2D_get_solution(PfemStep2D* step, double* p);
int32_t PfemStep2D_get_solution_morton(PfemStep2D* step, double* p);
/* nodal fields between input order and the Morton order of the step's last
 * mesh (permuteForward / permuteBackward, DirectAlphaShape2D.hpp:220-242);
 * host arrays of n elements of elem_bytes each */
int32_t PfemStep2D_permute_forward(PfemStep2D* step, const void* in, void* out, uint32_t elem_bytes);
int32_t PfemStep2D_permute_backward(PfemStep2D* step, const void* in, void* out, uint32_t elem_bytes);
int32_t PfemStep3D_permute_forward(PfemStep3D* step, const void* in, void* out, uint32_t elem_bytes);
int32_t PfemStep3D_permute_backward(PfemStep3D* step, const void* in, void* out, uint32_t elem_bytes);
uint32_t PfemStep2D_nnz(PfemStep2D* step);
int32_t PfemStep2D_get_system(PfemStep2D* step, uint32_t* row, uint32_t* col, double* val, double* b);
int32_t PfemStep2D_get_mesh(PfemStep2D* step, uint32_t* elem, uint8_t* node_is_bnd, double* xy, uint32_t* perm);
/* Average ms of a CG kernel on the last step's system (HIP events on the
 * library stream, `reps` launches): direction kernel of kind 0 = FP64
 * Jacobi path, 1 = FP32 row-sum Jacobi path, 2 = FP32 row-sum L L^T path;
 * kind 3 = the L L^T update kernel (with the solution update), 4 = the
 * L L^T lower apply, both on the learned factor (FSAI without a network). */
int32_t PfemStep2D_time_dir_kernel(PfemStep2D* step, int32_t kind, uint32_t reps, double* ms);
/* no reference counterpart: GPU ms of the mesh-and-assembly stages on the
 * current points, averaged over reps runs; ms [11] = box + Morton codes,
 * radix sort (all passes), gather, hierarchy, refit, candidate search, local
 * Delaunay + filter, scans, compress, P1 assembly, whole mesh + assembly */
int32_t PfemStep2D_time_sections(PfemStep2D* step, uint32_t reps, double* ms);

#ifdef __cplusplus
}
#endif
#endif
