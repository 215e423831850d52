// Fused device-resident CG kernels and host drivers (see numerics.hpp).
#include <cmath>
#include <cstring>
#include <vector>

#include "numerics.hpp"

namespace stream_amd {

constexpr int kCgThreads = 256;
// Grid of every CG kernel: at most 1024 blocks (4 per CU) with grid-stride
// rows, so the last block folds at most 4 partials per thread.
constexpr uint32_t kMaxCgBlocks = 1024;

Tuning& tuning() {
    static Tuning t;
    return t;
}

void CgWork::resize(uint32_t n) {
    blocks = (n + kCgThreads - 1) / kCgThreads;
    if (blocks > kMaxCgBlocks) blocks = kMaxCgBlocks;
    if (blocks == 0) blocks = 1;
    sc.resize(1);
    out2.resize(2);
    partials.resize(4ull * blocks);
    if (!counter.p) {
        counter.resize(1);
        fill<uint32_t>(counter.p, 1, 0u, ctx().stream);
    }
}

namespace {

template <int NT> __device__ __forceinline__ double block_sum(double v, double* s_tmp) {
    constexpr int NW = NT / kWave;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    const uint32_t w = threadIdx.x / kWave;
    if (lane_id() == 0) s_tmp[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < NW; ++k) t += s_tmp[k];
    __syncthreads();
    return t;  // valid in thread 0
}

// Fold K per-thread values over the grid.  Every block stores its K block
// sums; the last block to arrive sums all partials in block order and
// returns true in all its threads with the totals in tot[].  The ticket
// counter is reset by the last block.
template <int K>
__device__ bool grid_fold(const double (&v)[K], double* partials, uint32_t* counter, double (&tot)[K]) {
    __shared__ double s_tmp[kCgThreads / kWave];
    __shared__ double s_tot[K];
    __shared__ bool s_last;
    const uint32_t G = gridDim.x;  // <= kMaxCgBlocks
    double bs[K];
#pragma unroll
    for (int k = 0; k < K; ++k) bs[k] = block_sum<kCgThreads>(v[k], s_tmp);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) coherent_store(&partials[(size_t)k * G + blockIdx.x], bs[k]);
        vmem_release_order();
        const uint32_t t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (t == G - 1);
    }
    __syncthreads();
    if (!s_last) return false;
    vmem_acquire_order();
    constexpr int PER = kMaxCgBlocks / kCgThreads;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        // all loads issued before any add: one memory round trip
        double pv[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const uint32_t b = threadIdx.x + u * kCgThreads;
            pv[u] = b < G ? coherent_load(&partials[(size_t)k * G + b]) : 0.0;
        }
        double a = 0.0;
#pragma unroll
        for (int u = 0; u < PER; ++u) a += pv[u];
        const double t = block_sum<kCgThreads>(a, s_tmp);
        if (threadIdx.x == 0) s_tot[k] = t;
    }
    if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) tot[k] = s_tot[k];
    return true;
}

template <class MV, class XV>
__device__ __forceinline__ double row_dot(const uint32_t* __restrict__ row, const uint32_t* __restrict__ col,
                                          const MV* __restrict__ val, const XV* __restrict__ x, uint32_t i) {
    double s = 0.0;
    const uint32_t e = row[i + 1];
    for (uint32_t k = row[i]; k < e; ++k) s += (double)val[k] * (double)x[col[k]];
    return s;
}

#define CG_ROWS(i) for (uint32_t i = blockIdx.x * kCgThreads + threadIdx.x; i < n; i += gridDim.x * kCgThreads)

// ---------------------------------------------------------------------------
// init / replacement: r = b - A64 x64 (FP64), stored in the vector
// precision; Jacobi z; partials r.r and r.z (and b.b at init).
// mode 0: init (d_old = 0, beta = 0, iter = 0)
// mode 1: replacement of the recursive residual by the true one.  The
//         direction d and beta (computed by the recursion) are kept and only
//         r, z and r.z change: when the recursive residual had drifted far
//         below the true one, this degrades to a restart instead of scaling
//         the old direction by r_true.z / r_rec.z.
// ---------------------------------------------------------------------------
template <class VV, int PREC>
__global__ __launch_bounds__(kCgThreads) void cg_residual_kernel(
    uint32_t n, const uint32_t* __restrict__ row, const uint32_t* __restrict__ col, const double* __restrict__ val,
    const double* __restrict__ b, const double* __restrict__ x, VV* __restrict__ r, VV* __restrict__ z,
    VV* __restrict__ dold, const VV* __restrict__ idiag, CgScalars* sc, double* partials, uint32_t* counter,
    int mode, double rtol2) {
    double v[3] = {0.0, 0.0, 0.0};
    CG_ROWS(i) {
        const double ri = b[i] - row_dot(row, col, val, x, i);
        const VV rv = (VV)ri;
        r[i] = rv;
        if (mode == 0) dold[i] = (VV)0;
        v[0] += (double)rv * (double)rv;
        if constexpr (PREC != kPrecLLt) {
            const VV zv = (PREC == kPrecJacobi) ? (VV)(idiag[i] * rv) : rv;
            z[i] = zv;
            v[1] += (double)rv * (double)zv;
        }
        v[2] += b[i] * b[i];
    }
    double tot[3];
    if (!grid_fold<3>(v, partials, counter, tot)) return;
    if (threadIdx.x == 0) {
        sc->rr = tot[0];
        sc->aux = tot[0];
        if (mode == 0) {
            sc->iter = 0;
            sc->beta = 0.0;
            sc->dq = 0.0;
            sc->bb = tot[2];
            sc->tol2 = rtol2 >= 0.0 ? rtol2 * tot[2] : -rtol2;
        }
        if constexpr (PREC != kPrecLLt) {
            sc->rz = tot[1];
            if (mode == 0) sc->rz_old = tot[1];
        }
        const double cv = (sc->crit && PREC != kPrecLLt) ? tot[1] : tot[0];
        sc->done = (cv <= sc->tol2 || sc->iter >= sc->maxit) ? 1u : 0u;
    }
}

// d_new = z + beta d_old; q = A d_new; partial d_new.q.
// DIFF: the values are in row-sum form (off-diagonals K_ij, diagonal slot
// holding the row sum s_i) and q_i = sum_j K_ij (d_j - d_i) + s_i d_i.  For
// a graph Laplacian the differences of neighbouring entries of a smooth
// vector are exact in floating point (Sterbenz), so the FP32 product keeps
// its relative accuracy on the smooth modes that standard-form summation
// loses to cancellation.
template <class MV, class VV, bool DIFF>
__global__ __launch_bounds__(kCgThreads) void cg_dir_kernel(uint32_t n, const uint32_t* __restrict__ row,
                                                            const uint32_t* __restrict__ col,
                                                            const MV* __restrict__ val, const VV* __restrict__ z,
                                                            const VV* __restrict__ dold, VV* __restrict__ dnew,
                                                            VV* __restrict__ q, CgScalars* sc, double* partials,
                                                            uint32_t* counter) {
    if (sc->done) return;
    const VV beta = (VV)sc->beta;
    double v[1] = {0.0};
    CG_ROWS(i) {
        const VV dn = z[i] + beta * dold[i];
        dnew[i] = dn;
        VV s = (VV)0;
        const uint32_t e = row[i + 1];
        for (uint32_t k = row[i]; k < e; ++k) {
            const uint32_t j = col[k];
            if constexpr (DIFF) {
                if (j == i)
                    s += (VV)val[k] * dn;
                else
                    s += (VV)val[k] * ((z[j] + beta * dold[j]) - dn);
            } else {
                s += (VV)val[k] * (z[j] + beta * dold[j]);
            }
        }
        q[i] = s;
        v[0] += (double)dn * (double)s;
    }
    double tot[1];
    if (!grid_fold<1>(v, partials, counter, tot)) return;
    if (threadIdx.x == 0) {
        sc->dq = tot[0];
        sc->alpha = sc->rz / tot[0];
    }
}

// x += alpha d; r -= alpha q; z = M r (Jacobi / none); partial r.z, r.r
template <class VV, class XV, int PREC>
__global__ __launch_bounds__(kCgThreads) void cg_update_kernel(uint32_t n, XV* __restrict__ x, VV* __restrict__ r,
                                                               VV* __restrict__ z, const VV* __restrict__ d,
                                                               const VV* __restrict__ q, const VV* __restrict__ idiag,
                                                               CgScalars* sc, double* partials, uint32_t* counter) {
    if (sc->done) return;
    const double alpha = sc->alpha;
    const VV av = (VV)alpha;
    double v[2] = {0.0, 0.0};
    CG_ROWS(i) {
        x[i] += (XV)(alpha * (double)d[i]);
        const VV rv = r[i] - av * q[i];
        r[i] = rv;
        v[1] += (double)rv * (double)rv;
        if constexpr (PREC != kPrecLLt) {
            const VV zv = (PREC == kPrecJacobi) ? idiag[i] * rv : rv;
            z[i] = zv;
            v[0] += (double)rv * (double)zv;
        }
    }
    double tot[2];
    if (!grid_fold<2>(v, partials, counter, tot)) return;
    if (threadIdx.x == 0) {
        sc->rr = tot[1];
        if constexpr (PREC != kPrecLLt) {
            sc->rz_old = sc->rz;
            sc->beta = tot[0] / sc->rz;
            sc->rz = tot[0];
            sc->iter += 1;
            const double cv = sc->crit ? tot[0] : tot[1];
            sc->done = (cv <= sc->tol2 || sc->iter >= sc->maxit) ? 1u : 0u;
        }
    }
}

// w = L^T r : w_i = L_ii r_i + sum_{j>i} L_ji r_j
template <class VV>
__global__ __launch_bounds__(kCgThreads) void llt_up_kernel(uint32_t n, const uint32_t* __restrict__ up_row,
                                                            const uint32_t* __restrict__ up_col,
                                                            const VV* __restrict__ up_val, const VV* __restrict__ diag,
                                                            const VV* __restrict__ r, VV* __restrict__ w,
                                                            const CgScalars* sc, int force) {
    if (!force && sc->done) return;
    CG_ROWS(i) {
        VV s = diag[i] * r[i];
        const uint32_t e = up_row[i + 1];
        for (uint32_t k = up_row[i]; k < e; ++k) s += up_val[k] * r[up_col[k]];
        w[i] = s;
    }
}

// z = L w : z_i = L_ii w_i + sum_{j<i} L_ij w_j; partial r.z.
// mode 0: iteration bookkeeping (beta, iter, done); mode 1: init;
// mode 2: after a residual replacement (beta kept, as in cg_residual_kernel).
template <class VV>
__global__ __launch_bounds__(kCgThreads) void llt_lo_kernel(uint32_t n, const uint32_t* __restrict__ lo_row,
                                                            const uint32_t* __restrict__ lo_col,
                                                            const VV* __restrict__ lo_val, const VV* __restrict__ diag,
                                                            const VV* __restrict__ w, const VV* __restrict__ r,
                                                            VV* __restrict__ z, CgScalars* sc, double* partials,
                                                            uint32_t* counter, int mode) {
    if (mode == 0 && sc->done) return;
    double v[1] = {0.0};
    CG_ROWS(i) {
        VV s = diag[i] * w[i];
        const uint32_t e = lo_row[i + 1];
        for (uint32_t k = lo_row[i]; k < e; ++k) s += lo_val[k] * w[lo_col[k]];
        z[i] = s;
        v[0] += (double)r[i] * (double)s;
    }
    double tot[1];
    if (!grid_fold<1>(v, partials, counter, tot)) return;
    if (threadIdx.x == 0) {
        if (mode == 0) {
            sc->rz_old = sc->rz;
            sc->beta = tot[0] / sc->rz;
            sc->rz = tot[0];
            sc->iter += 1;
            const double cv = sc->crit ? tot[0] : sc->rr;
            sc->done = (cv <= sc->tol2 || sc->iter >= sc->maxit) ? 1u : 0u;
        } else if (mode == 1) {
            sc->rz = tot[0];
            sc->rz_old = tot[0];
        } else {
            sc->rz = tot[0];
        }
    }
}

// x64 += delta32; delta32 = 0
__global__ __launch_bounds__(kCgThreads) void merge_delta_kernel(uint32_t n, double* __restrict__ x,
                                                                 float* __restrict__ delta) {
    CG_ROWS(i) {
        x[i] += (double)delta[i];
        delta[i] = 0.0f;
    }
}

__global__ __launch_bounds__(kCgThreads) void true_residual_kernel(uint32_t n, const uint32_t* __restrict__ row,
                                                                   const uint32_t* __restrict__ col,
                                                                   const double* __restrict__ val,
                                                                   const double* __restrict__ b,
                                                                   const double* __restrict__ x, double* partials,
                                                                   uint32_t* counter, double* out2) {
    double v[2] = {0.0, 0.0};
    CG_ROWS(i) {
        const double ri = b[i] - row_dot(row, col, val, x, i);
        v[0] += ri * ri;
        v[1] += b[i] * b[i];
    }
    double tot[2];
    if (!grid_fold<2>(v, partials, counter, tot)) return;
    if (threadIdx.x == 0) {
        out2[0] = tot[0];
        out2[1] = tot[1];
    }
}

CgScalars read_sc(CgWork& w, hipStream_t s) {
    CgScalars h;
    SA_HIP(hipMemcpyAsync(ctx().pinned, w.sc.p, sizeof(CgScalars), hipMemcpyDeviceToHost, s));
    SA_HIP(hipStreamSynchronize(s));
    std::memcpy(&h, ctx().pinned, sizeof(h));
    return h;
}

void publish_options(CgWork& w, const CgOptions& o, hipStream_t s) {
    // maxit and the criterion are written before the init pass; everything
    // else is produced by the kernels
    CgScalars h{};
    h.maxit = o.maxit;
    h.crit = o.crit;
    SA_HIP(hipMemcpyAsync(w.sc.p, &h, sizeof(h), hipMemcpyHostToDevice, s));
}

void clear_done(CgWork& w, hipStream_t s) {
    static const uint32_t zero = 0;
    SA_HIP(hipMemcpyAsync(&w.sc.p->done, &zero, sizeof(uint32_t), hipMemcpyHostToDevice, s));
    SA_HIP(hipStreamSynchronize(s));
}

}  // namespace

void true_residual(const DevCsr<double>& A, const double* b, const double* x, double* out2, CgWork& w,
                   hipStream_t s) {
    w.resize(A.n);
    hipLaunchKernelGGL(true_residual_kernel, dim3(w.blocks), dim3(kCgThreads), 0, s, A.n, A.row.p, A.col.p, A.val.p,
                       b, x, w.partials.p, w.counter.p, w.out2.p);
    d2h(out2, w.out2.p, 2);
}

// ---------------------------------------------------------------------------
// Generic driver.  VV: vector precision, MV: matrix precision of the
// iteration, XV: accumulator of the solution updates (double: x itself;
// float: delta merged into x64 at replacements).
// ---------------------------------------------------------------------------
template <class VV, class MV, class XV, int PREC, bool DIFF = false>
static void cg_drive(const DevCsr<double>& A64, const uint32_t* row, const uint32_t* col, const MV* val,
                     const VV* idiag, const LowerFactor<VV>* L, const double* b, double* x, const CgOptions& o,
                     CgWork& w, SolveStats& st, hipStream_t s) {
    const uint32_t n = A64.n;
    w.resize(n);
    const dim3 grid(w.blocks), blk(kCgThreads);
    // vectors carved from the grow-only workspace: r z d0 d1 q w delta
    const size_t nv = ((size_t)n + 63) & ~(size_t)63;
    w.vbuf.resize(7 * nv * sizeof(double));
    VV* base = reinterpret_cast<VV*>(w.vbuf.p);
    VV* r = base;
    VV* z = base + nv;
    VV* d0 = base + 2 * nv;
    VV* d1 = base + 3 * nv;
    VV* q = base + 4 * nv;
    VV* wv = base + 5 * nv;
    float* delta = reinterpret_cast<float*>(base + 6 * nv);
    XV* xacc;
    if constexpr (sizeof(XV) == 4) {
        fill<float>(delta, n, 0.0f, s);
        xacc = (XV*)delta;
    } else {
        xacc = (XV*)x;
    }
    hipEvent_t e0, e1;
    SA_HIP(hipEventCreate(&e0));
    SA_HIP(hipEventCreate(&e1));
    SA_HIP(hipEventRecord(e0, s));

    publish_options(w, o, s);
    const double rtol2 = o.abs_tol2 >= 0.0 ? -o.abs_tol2 : o.rtol * o.rtol;
    auto residual = [&](int mode) {
        hipLaunchKernelGGL((cg_residual_kernel<VV, PREC>), grid, blk, 0, s, n, A64.row.p, A64.col.p, A64.val.p, b,
                           x, r, z, d0, idiag, w.sc.p, w.partials.p, w.counter.p, mode, rtol2);
        if constexpr (PREC == kPrecLLt) {
            hipLaunchKernelGGL((llt_up_kernel<VV>), grid, blk, 0, s, n, L->up_row.p, L->up_col.p, L->up_val.p,
                               L->diag.p, r, wv, w.sc.p, 1);
            hipLaunchKernelGGL((llt_lo_kernel<VV>), grid, blk, 0, s, n, L->lo_row.p, L->lo_col.p, L->lo_val.p,
                               L->diag.p, wv, r, z, w.sc.p, w.partials.p, w.counter.p, mode == 0 ? 1 : 2);
        }
    };
    auto merge = [&] {
        if constexpr (sizeof(XV) == 4) hipLaunchKernelGGL(merge_delta_kernel, grid, blk, 0, s, n, x, delta);
    };
    residual(0);
    CgScalars h = read_sc(w, s);
    if (o.min_one && h.done) {
        clear_done(w, s);
        h.done = 0;
    }
    double rr_last_rep = h.rr;
    uint32_t reps = 0;
    while (true) {
        if (h.done) {
            if (h.iter >= o.maxit || !o.verify) break;
            // the recursive residual says converged: verify with the FP64
            // true residual, which also replaces the recursive one
            merge();
            residual(1);
            ++reps;
            h = read_sc(w, s);
            rr_last_rep = h.aux;
            if (h.aux <= h.tol2 || h.iter >= o.maxit) break;
            continue;
        }
        const uint32_t it0 = h.iter;
        for (uint32_t k = 0; k < o.batch; ++k) {
            const uint32_t it = it0 + k;
            VV* dold = (it & 1) ? d1 : d0;
            VV* dnew = (it & 1) ? d0 : d1;
            hipLaunchKernelGGL((cg_dir_kernel<MV, VV, DIFF>), grid, blk, 0, s, n, row, col, val, z, dold, dnew, q, w.sc.p,
                               w.partials.p, w.counter.p);
            hipLaunchKernelGGL((cg_update_kernel<VV, XV, PREC>), grid, blk, 0, s, n, xacc, r, z, dnew, q, idiag,
                               w.sc.p, w.partials.p, w.counter.p);
            if constexpr (PREC == kPrecLLt) {
                hipLaunchKernelGGL((llt_up_kernel<VV>), grid, blk, 0, s, n, L->up_row.p, L->up_col.p, L->up_val.p,
                                   L->diag.p, r, wv, w.sc.p, 0);
                hipLaunchKernelGGL((llt_lo_kernel<VV>), grid, blk, 0, s, n, L->lo_row.p, L->lo_col.p, L->lo_val.p,
                                   L->diag.p, wv, r, z, w.sc.p, w.partials.p, w.counter.p, 0);
            }
        }
        h = read_sc(w, s);
        if (!h.done && o.replace_factor > 0 && h.rr < o.replace_factor * o.replace_factor * rr_last_rep) {
            merge();
            residual(1);
            ++reps;
            h = read_sc(w, s);
            rr_last_rep = h.aux;
        }
    }
    merge();
    double tr[2];
    hipLaunchKernelGGL(true_residual_kernel, grid, blk, 0, s, n, A64.row.p, A64.col.p, A64.val.p, b, x, w.partials.p,
                       w.counter.p, w.out2.p);
    d2h(tr, w.out2.p, 2);
    SA_HIP(hipEventRecord(e1, s));
    SA_HIP(hipEventSynchronize(e1));
    float ms = 0;
    SA_HIP(hipEventElapsedTime(&ms, e0, e1));
    SA_HIP(hipEventDestroy(e0));
    SA_HIP(hipEventDestroy(e1));
    h = read_sc(w, s);
    st.iters = h.iter;
    st.replacements = reps;
    st.true_rel = tr[1] > 0 ? std::sqrt(tr[0] / tr[1]) : std::sqrt(tr[0]);
    st.ms = ms;
}

void cg_fp64(const DevCsr<double>& A, const double* idiag, const double* b, double* x, double rtol, uint32_t maxit,
             CgWork& w, SolveStats& st, hipStream_t s) {
    CgOptions o;
    o.rtol = rtol;
    o.maxit = maxit;
    cg_drive<double, double, double, kPrecJacobi>(A, A.row.p, A.col.p, A.val.p, idiag, nullptr, b, x, o, w, st, s);
}

void cg_mixed(const DevCsr<double>& A64, const DevCsr<float>& A32, const float* idiag32, const LowerFactor<float>* L,
              int32_t prec, const double* b, double* x, double rtol, uint32_t maxit, CgWork& w, SolveStats& st,
              hipStream_t s) {
    CgOptions o;
    o.rtol = rtol;
    o.maxit = maxit;
    o.replace_factor = tuning().replace_factor;
    o.batch = tuning().batch;
    if (prec == kPrecLLt) {
        if (!L) throw std::invalid_argument("L L^T preconditioner requested without a factor");
        cg_drive<float, float, float, kPrecLLt, true>(A64, A64.row.p, A64.col.p, A32.val.p, nullptr, L, b, x, o, w, st,
                                                      s);
    } else {
        cg_drive<float, float, float, kPrecJacobi, true>(A64, A64.row.p, A64.col.p, A32.val.p, idiag32, nullptr, b, x,
                                                         o, w, st, s);
    }
}

void cg_fp32_opts(const DevCsr<double>& A64, const float* A32val, const float* idiag32, int32_t prec,
                  const double* b, double* x, const CgOptions& o, CgWork& w, SolveStats& st, hipStream_t s,
                  bool rowsum_form) {
    if (rowsum_form) {
        if (prec != kPrecJacobi) throw std::invalid_argument("row-sum form is used with the Jacobi preconditioner");
        cg_drive<float, float, float, kPrecJacobi, true>(A64, A64.row.p, A64.col.p, A32val, idiag32, nullptr, b, x, o,
                                                         w, st, s);
    } else if (prec == kPrecJacobi) {
        cg_drive<float, float, float, kPrecJacobi>(A64, A64.row.p, A64.col.p, A32val, idiag32, nullptr, b, x, o, w,
                                                   st, s);
    } else {
        cg_drive<float, float, float, kPrecNone>(A64, A64.row.p, A64.col.p, A32val, nullptr, nullptr, b, x, o, w, st,
                                                 s);
    }
}

// Average duration (ms, HIP events on the context stream) of the CG
// direction kernel (d update + SpMV + d.q) over `reps` back-to-back
// launches on the given system, FP64 or FP32 values.  The solver state is
// scratch: the vectors are filled with ones first.
double time_dir_kernel(const DevCsr<double>& A, const float* val32, bool fp32, uint32_t reps, CgWork& w,
                       hipStream_t s) {
    const uint32_t n = A.n;
    w.resize(n);
    const size_t nv = ((size_t)n + 63) & ~(size_t)63;
    w.vbuf.resize(7 * nv * sizeof(double));
    CgScalars h{};
    h.maxit = 0xffffffffu;
    h.rz = 1.0;
    SA_HIP(hipMemcpyAsync(w.sc.p, &h, sizeof(h), hipMemcpyHostToDevice, s));
    const dim3 grid(w.blocks), blk(kCgThreads);
    hipEvent_t e0, e1;
    SA_HIP(hipEventCreate(&e0));
    SA_HIP(hipEventCreate(&e1));
    auto launch = [&](uint32_t k) {
        if (fp32) {
            float* base = reinterpret_cast<float*>(w.vbuf.p);
            float *z = base, *d0 = base + nv, *d1 = base + 2 * nv, *q = base + 3 * nv;
            hipLaunchKernelGGL((cg_dir_kernel<float, float, true>), grid, blk, 0, s, n, A.row.p, A.col.p, val32, z,
                               (k & 1) ? d1 : d0, (k & 1) ? d0 : d1, q, w.sc.p, w.partials.p, w.counter.p);
        } else {
            double* base = reinterpret_cast<double*>(w.vbuf.p);
            double *z = base, *d0 = base + nv, *d1 = base + 2 * nv, *q = base + 3 * nv;
            hipLaunchKernelGGL((cg_dir_kernel<double, double, false>), grid, blk, 0, s, n, A.row.p, A.col.p, A.val.p, z,
                               (k & 1) ? d1 : d0, (k & 1) ? d0 : d1, q, w.sc.p, w.partials.p, w.counter.p);
        }
    };
    if (fp32)
        fill<float>(reinterpret_cast<float*>(w.vbuf.p), 4 * nv, 1.0f, s);
    else
        fill<double>(reinterpret_cast<double*>(w.vbuf.p), 4 * nv, 1.0, s);
    for (uint32_t k = 0; k < 3; ++k) launch(k);
    SA_HIP(hipEventRecord(e0, s));
    for (uint32_t k = 0; k < reps; ++k) launch(k);
    SA_HIP(hipEventRecord(e1, s));
    SA_HIP(hipEventSynchronize(e1));
    float ms = 0;
    SA_HIP(hipEventElapsedTime(&ms, e0, e1));
    SA_HIP(hipEventDestroy(e0));
    SA_HIP(hipEventDestroy(e1));
    return (double)ms / reps;
}

}  // namespace stream_amd
