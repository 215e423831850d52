// Sparse linear algebra for the pressure solve.
//
// Replaces src/numerics (DeviceCSR/HostCSR csr.cpp:26-101, LinearSystem
// linear_system.cpp:27-43, PrecJacobi preconditioner_jacobi.cpp:29-60,
// ConjugateGradient conjugate_gradient.cpp:44-208).
//
// Matrix layout for the iterations: packed SELL-64.  The off-diagonal
// entries of 64 consecutive rows (one wavefront, one row per lane) form a
// slice stored step by step: step k holds the k-th entries of the rows
// longer than k, in lane order and without gaps, so the k-th load of a
// wave reads one contiguous run and no slot is padding (a slice of rows of
// 3 and 6 entries costs 3 + 6 slots, not 2 x 6).  Entry k of row r (slice
// s = r/64, lane l = r%64) lives at off[s] + sum over the slice's rows of
// min(len, k) + (number of lanes below l whose row is longer than k); a
// wave computes that from a ballot of (len > k) and a lane-mask count, so
// the row loops load no index table.  The slice width (longest row) bounds
// the steps.  The diagonal is a separate array.  CSR stays the interchange
// format (reference DeviceCSR, assembly output, FP64 true residuals).
//
// CG keeps every scalar on the device.  One iteration is two kernels:
//   dir    : d = z + beta d_old on the own row, q = A d with neighbour
//            entries of d recomputed from the interleaved {z, d_old} pair
//            (one load per neighbour), partial d.q;
//   update : x += alpha d, r -= alpha q, z = M^-1 r, partial r.z and r.r.
// The two vectors z and d live interleaved in two ping-pong buffers:
// iteration k reads P[k&1] = {z_k, d_{k-1}} and writes P[(k+1)&1] =
// {z_{k+1}, d_k}.  Each kernel folds its block partials with a
// last-block-done reduction, so alpha, beta and the convergence flag never
// leave the device; the host reads them once per batch of iterations.
// Dots accumulate in FP64 whatever the vector precision.
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

constexpr uint32_t kSlice = 64;

// SELL column offsets: entries whose column is more than 32767 rows away
// from their row hold kColEscape in col16 and are read from col.
constexpr int16_t kColEscape = -32768;

template <class V> struct Sell {
    uint32_t n = 0;
    uint32_t nslices = 0;
    uint32_t entries = 0;    // slots (= off-diagonal entries; the arrays hold kSlice more)
    DBuf<uint32_t> off;      // [nslices + 1] first slot of each slice
    DBuf<uint16_t> width;    // [nslices]
    DBuf<uint16_t> len;      // [n]
    DBuf<uint32_t> col;      // [entries]
    DBuf<int16_t> col16;     // [entries] column - row, or kColEscape where that does not fit
    DBuf<V> val;             // [entries]
    DBuf<V> diag;            // [n]
    DBuf<uint32_t> tmp;      // build scratch
    DBuf<uint64_t> scan_ws;
};

// Value forms of a SELL copy.
enum SellForm : int32_t {
    kSellStandard = 0,  // off-diagonals and diagonal as in the CSR
    kSellRowSum = 1,    // off-diagonals; diagonal slot = FP64 row sum (difference form)
};
enum SellPart : int32_t { kSellOffDiag = 0, kSellLower = 1, kSellUpper = 2 };

// Build a SELL copy of a CSR matrix (values from val64 or val32, one of them
// non-null).  part selects all off-diagonals, the strict lower or the strict
// upper triangle.  entry_slot (optional, [nnz]): entry_slot[e] = the slot of
// CSR entry e, written for the entries of the part only.  Synchronises once
// (allocation size).
template <class V>
void build_sell(uint32_t n, const uint32_t* row, const uint32_t* col, const double* val64, const float* val32,
                SellForm form, SellPart part, Sell<V>& S, hipStream_t s, uint32_t* entry_slot = nullptr);

// Device-resident CG scalars (one 128-byte record).
struct CgScalars {
    double rz;      // r.z of the current iterate
    double rz_old;  // previous r.z
    double dq;      // d.q
    double rr;      // r.r (recursive residual)
    double alpha;
    double beta;
    double tol2;    // stop when the criterion <= tol2 (= rtol^2 * bb)
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
    uint32_t threads = 256;  // block size of the last resize
    // grid for n rows with blocks of nthreads (256, 512 or 1024)
    void resize(uint32_t n, uint32_t nthreads = 256);
};

// Preconditioner kind of the fused solvers.
enum PrecKind : int32_t { kPrecNone = 0, kPrecJacobi = 1, kPrecLLt = 2 };

// Factor of M^-1 = L L^T, L lower triangular on A's pattern: lo holds the
// strictly lower entries L_ij (j < i) by row, up the strictly upper entries
// of row i holding L_ji (the strict upper part of L^T), diag L_ii.
template <class V> struct LowerFactor {
    uint32_t n = 0;
    Sell<V> lo, up;
    DBuf<V> diag;
    // per CSR entry of the pattern the halves were built from, its slot in
    // lo (column < row) or up (column > row)
    DBuf<uint32_t> slot;
};

struct SolveStats {
    uint32_t iters = 0;        // iterations
    uint32_t replacements = 0; // FP64 residual replacements
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
    double replace_factor = 0.0; // replace when ||r|| fell by this factor
    uint32_t batch = 32;         // iterations between host checks
};

// Process-wide solver knobs (set through Stream_tune).
struct Tuning {
    double replace_factor = 1e-1;  // mixed: residual replacement when ||r|| fell by this factor
    uint32_t batch = 32;           // CG iterations between host checks
    uint32_t cg_blocks = 2048;     // grid cap of the CG kernels in 256-thread blocks (64..4096)
    uint32_t graph = 0;            // 1: replay each batch of iterations as a captured HIP graph
    uint32_t xcd_map = 1;          // SELL slices: 0 grid stride, 1 XCD-contiguous chunks per wave
    uint32_t fold_group = 32;      // blocks per group of the two-level grid fold (0: one level)
    uint32_t sell_chunk = 8;       // SELL entries per load step of the CG row loops (4 or 8)
    uint32_t llt_threads = 512;    // block size of the L L^T path's kernels (256, 512 or 1024)
};
Tuning& tuning();
// Group size of the two-level grid fold of the CG kernels (0, 8, 16, 32, 64).
void set_fold_group(uint32_t g);

// CG on SELL matrix S (values V) with vectors in precision V.  A64 (CSR,
// FP64) gives the true residuals.  x: FP64 solution (in/out), updated in
// FP64 every iteration.  diff: S holds the row-sum form.
template <class V>
void cg_solve(const DevCsr<double>& A64, const Sell<V>& S, bool diff, int32_t prec, const V* idiag,
              const LowerFactor<V>* L, const double* b, double* x, const CgOptions& o, CgWork& w, SolveStats& st,
              hipStream_t s);

// ||b - A x||_2^2 and ||b||_2^2 in FP64.
void true_residual(const DevCsr<double>& A, const double* b, const double* x, double* out2, CgWork& w, hipStream_t s);

// Average duration (ms) of a CG direction kernel over `reps` launches (llt:
// the L L^T path's kernel, else the Jacobi path's).
template <class V>
double time_dir_kernel(const Sell<V>& S, bool diff, bool llt, uint32_t reps, CgWork& w, hipStream_t s);
// Average duration (ms) of the L L^T update kernel (lower: the lower apply)
// over `reps` launches on L.
template <class V>
double time_llt_update_kernel(const LowerFactor<V>& L, uint32_t reps, CgWork& w, hipStream_t s, bool lower);

}  // namespace stream_amd
