// Sparse linear algebra for the pressure solve.
//
// Replaces src/numerics (DeviceCSR/HostCSR csr.cpp:26-101, LinearSystem
// linear_system.cpp:27-43, PrecJacobi preconditioner_jacobi.cpp:29-60,
// ConjugateGradient conjugate_gradient.cpp:44-208).
//
// CG on the GPU keeps every scalar on the device.  One iteration is two
// kernels:
//   dir    : d_new = z + beta*d_old for the own rows, q = A d_new with the
//            neighbour entries of d_new recomputed from (z, d_old), and the
//            partial sums of d_new.q;
//   update : x += alpha d, r -= alpha q, z = M^-1 r, partial r.z and r.r.
// Each kernel folds its block partials with a last-block-done reduction
// (partials exchanged through coherent stores), so the scalars alpha, beta
// and the convergence flag are produced on the device; the host reads them
// once per batch of iterations.  Dots accumulate in FP64 whatever the
// vector precision.
#pragma once
#include "geometry.hpp"
#include "primitives.hpp"

namespace stream_amd {

template <class V> struct DevCsr {
    uint32_t n = 0;
    uint32_t nnz = 0;
    DBuf<uint32_t> row, col;
    DBuf<V> val;
};

// Device-resident CG scalars (one 128-byte record).
struct CgScalars {
    double rz;      // r.z of the current iterate
    double rz_old;  // previous r.z
    double dq;      // d.q
    double rr;      // r.r (recursive residual)
    double alpha;
    double beta;
    double tol2;    // stop when rr <= tol2 (= rtol^2 * bb)
    double aux;     // r.r of the last true-residual (replacement) pass
    double bb;      // b.b
    uint32_t iter;
    uint32_t maxit;
    uint32_t done;
    uint32_t crit;  // 0: stop on r.r <= tol2, 1: stop on r.z <= tol2
    uint32_t pad[2];
};

struct CgWork {
    DBuf<CgScalars> sc;
    DBuf<double> partials;   // [3][blocks]
    DBuf<uint32_t> counter;  // last-block tickets
    DBuf<double> out2;       // true-residual norms
    DBuf<uint8_t> vbuf;      // solver vectors (grow-only)
    uint32_t blocks = 0;
    void resize(uint32_t n);
};

// Preconditioner kind of the fused solvers.
enum PrecKind : int32_t { kPrecNone = 0, kPrecJacobi = 1, kPrecLLt = 2 };

// Learned factor: M^-1 = L L^T with L lower triangular on A's pattern.
// lo: strictly lower entries (j < i) of each row, up: strictly upper entries
// (j > i) holding L_ji, diag: L_ii.  Both halves in CSR of their own.
template <class V> struct LowerFactor {
    uint32_t n = 0;
    DBuf<uint32_t> lo_row, lo_col, up_row, up_col;
    DBuf<V> lo_val, up_val, diag;
};

struct SolveStats {
    uint32_t iters = 0;        // inner iterations
    uint32_t replacements = 0; // FP64 residual replacements (mixed)
    double true_rel = 0.0;     // ||b - A x|| / ||b|| in FP64
    double ms = 0.0;
};

struct CgOptions {
    double rtol = 1e-8;          // relative: tol2 = rtol^2 ||b||^2 (when abs_tol2 < 0)
    double abs_tol2 = -1.0;      // absolute threshold on the criterion when >= 0
    uint32_t maxit = 100000;
    uint32_t crit = 0;           // 0: r.r, 1: r.z
    bool verify = true;          // FP64 true-residual verification at convergence
    bool min_one = false;        // run at least one iteration (reference loop)
    double replace_factor = 0.0; // mixed: replace when ||r|| fell by this factor
    uint32_t batch = 32;         // iterations between host checks
};

// Process-wide solver knobs (set through Stream_tune).
struct Tuning {
    double replace_factor = 1e-1;  // mixed: residual replacement when ||r|| fell by this factor
    uint32_t batch = 32;           // CG iterations between host checks
};
Tuning& tuning();

// FP64 Jacobi-PCG to ||b - Ax|| / ||b|| <= rtol (true residual verified;
// restarts from the true residual when the recursive one drifted).
void cg_fp64(const DevCsr<double>& A, const double* idiag, const double* b, double* x, double rtol, uint32_t maxit,
             CgWork& w, SolveStats& st, hipStream_t s);
// The same driver with FP32 vectors and matrix values (A32val on A's
// pattern) under explicit options.  rowsum_form: A32val holds the
// off-diagonals and, in the diagonal slot, the row sum (Laplacian
// difference form of the product).
void cg_fp32_opts(const DevCsr<double>& A64, const float* A32val, const float* idiag32, int32_t prec,
                  const double* b, double* x, const CgOptions& o, CgWork& w, SolveStats& st, hipStream_t s,
                  bool rowsum_form = false);

// Mixed precision: FP32 matrix, FP32 vectors and preconditioner, FP64 dots,
// x accumulated in FP64 and the residual replaced by the FP64 true residual
// b - A x whenever the recursive residual has dropped by `replace_factor`
// since the last replacement (and at convergence).  prec: kPrecJacobi uses
// idiag32; kPrecLLt uses L.
void cg_mixed(const DevCsr<double>& A64, const DevCsr<float>& A32, const float* idiag32, const LowerFactor<float>* L,
              int32_t prec, const double* b, double* x, double rtol, uint32_t maxit, CgWork& w, SolveStats& st,
              hipStream_t s);

// ||b - A x||_2 and ||b||_2 in FP64.
void true_residual(const DevCsr<double>& A, const double* b, const double* x, double* out2, CgWork& w, hipStream_t s);

// Average duration (ms) of the CG direction kernel over `reps` launches.
double time_dir_kernel(const DevCsr<double>& A, const float* val32, bool fp32, uint32_t reps, CgWork& w,
                       hipStream_t s);

}  // namespace stream_amd
