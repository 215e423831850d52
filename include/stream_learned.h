/*
 * Learned preconditioner C-ABI of the MI355X build (libstream_amd.so).
 *
 * No counterpart in the reference: its only preconditioner is PrecJacobi
 * (include/numerics/preconditioner.hpp:60-72, created and applied through
 * python/numerics/jacobi.py), so this handle follows PrecJacobi's
 * create/destroy pattern and plugs into the step driver as solver 3.
 *
 * A message-passing network (3 layers, hidden 64, BF16 weights on MFMA with
 * FP32 accumulation) emits the values of an upper-triangular Ghat on the
 * pattern of A; the preconditioner is M^-1 = L L^T with L = D^-1/2 Ghat^T
 * (D = diag A), SPD by construction.  Training minimises Kaporin's
 * condition number of Ghat (D^-1/2 A D^-1/2) Ghat^T with Adam on the GPU.
 * Parameters are one FP32 array of PrecLearned_param_count() values.
 */
#ifndef STREAM_AMD_LEARNED_H
#define STREAM_AMD_LEARNED_H
#include <stdint.h>

#include "stream_numerics.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct PrecLearned PrecLearned;

PrecLearned* PrecLearned_create(void);
void PrecLearned_destroy(PrecLearned* net);
uint32_t PrecLearned_param_count(void);
/* FP32 master parameters; set_params also resets the Adam state. */
int32_t PrecLearned_set_params(PrecLearned* net, const float* w, uint32_t count);
int32_t PrecLearned_get_params(PrecLearned* net, float* w, uint32_t count);
int32_t PrecLearned_set_adam(PrecLearned* net, float beta1, float beta2, float eps);
/* Host system: CSR (sorted columns, symmetric pattern, FP64 values), points
 * [n][2] in row order, Dirichlet flags [n].  Copied to the device.  A row
 * whose columns are not strictly ascending is refused (invalid argument). */
int32_t PrecLearned_set_system(PrecLearned* net, uint32_t n, const uint32_t* row, const uint32_t* col,
                               const double* val, const double* xy, const uint8_t* bnd);
/* the same with coordinates [n][dim], dim 2 or 3 */
int32_t PrecLearned_set_system_dim(PrecLearned* net, uint32_t dim, uint32_t n, const uint32_t* row,
                                   const uint32_t* col, const double* val, const double* xyz, const uint8_t* bnd);
/* The step's current (assembled) system, device-resident. */
int32_t PrecLearned_bind_step(PrecLearned* net, PfemStep2D* step);
/* The same for the 3D step (node features use x, y, z). */
int32_t PrecLearned_bind_step3d(PrecLearned* net, PfemStep3D* step);
/* Network of the 3D step's solver 3 (not owned; NULL clears). */
int32_t PfemStep3D_set_learned(PfemStep3D* step, PrecLearned* net);
/* Forward + loss; ghat [nnz] (nullable) receives Ghat per CSR entry (upper
 * entries and diagonal). */
int32_t PrecLearned_forward(PrecLearned* net, double* loss, float* ghat);
/* Forward + loss + gradient w.r.t. the FP32 parameters (grad [count]). */
int32_t PrecLearned_gradient(PrecLearned* net, double* loss, float* grad);
/* One Adam step with learning rate lr; loss (nullable) before the update. */
int32_t PrecLearned_train_step(PrecLearned* net, float lr, double* loss);
/* Average ms over `reps` launches (HIP events) of the layer-0 MFMA GEMM
 * [P|Q] = H W_pq on the bound system; flops (nullable) = 2 n 64 128. */
int32_t PrecLearned_time_gemm(PrecLearned* net, uint32_t reps, double* ms, double* flops);
/* Network used by PfemStep2D_run's solver 3 (not owned; NULL detaches). */
int32_t PfemStep2D_set_learned(PfemStep2D* step, PrecLearned* net);

#ifdef __cplusplus
}
#endif
#endif
