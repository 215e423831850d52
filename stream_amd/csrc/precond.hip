// FSAI factor and L L^T structure (see precond.hpp).
#include <cmath>

#include "precond.hpp"

namespace stream_amd {

namespace {

constexpr uint32_t kFsaiRegs = 8;  // dense solve held in registers up to 8 x 8

// Row i of the upper factor G: P = {i, upper neighbours...} (at most 8,
// the first ones in column order when the row is longer), A[P,P] g = e_0
// by Cholesky in FP64, g scaled by 1/sqrt(g_0).
template <class V>
__global__ __launch_bounds__(256) void fsai_kernel(uint32_t n, const uint32_t* __restrict__ row,
                                                   const uint32_t* __restrict__ col, const double* __restrict__ val,
                                                   const uint16_t* __restrict__ up_len,
                                                   const uint32_t* __restrict__ slot, V* __restrict__ up_val,
                                                   V* __restrict__ lo_val, V* __restrict__ diag) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t lu = up_len[i];
    const uint32_t m = lu + 1 < kFsaiRegs ? lu + 1 : kFsaiRegs;
    // P: i, then the first upper entries of row i in CSR order (the order of
    // the factor's upper slots)
    uint32_t P[kFsaiRegs];
    P[0] = i;
#pragma unroll
    for (uint32_t k = 1; k < kFsaiRegs; ++k) P[k] = 0xffffffffu;
    {
        uint32_t k = 1;
        for (uint32_t e = row[i]; e < row[i + 1] && k < m; ++e)
            if (col[e] > i) P[k++] = col[e];
    }
    // M[b][a] (a <= b) packed lower triangle, row-major
    double M[kFsaiRegs * (kFsaiRegs + 1) / 2];
#pragma unroll
    for (uint32_t t = 0; t < kFsaiRegs * (kFsaiRegs + 1) / 2; ++t) M[t] = 0.0;
#pragma unroll
    for (uint32_t b = 0; b < kFsaiRegs; ++b) {
        if (b < m) {
            const uint32_t rb = P[b];
            for (uint32_t e = row[rb]; e < row[rb + 1]; ++e) {
                const uint32_t c = col[e];
                const double v = val[e];
#pragma unroll
                for (uint32_t a = 0; a <= b; ++a)
                    if (P[a] == c) M[b * (b + 1) / 2 + a] = v;
            }
        }
    }
    // Cholesky M = C C^T (in place); rows/cols >= m stay identity
#pragma unroll
    for (uint32_t b = 0; b < kFsaiRegs; ++b) {
        if (b >= m) {
            M[b * (b + 1) / 2 + b] = 1.0;
            continue;
        }
#pragma unroll
        for (uint32_t a = 0; a <= b; ++a) {
            double sum = M[b * (b + 1) / 2 + a];
#pragma unroll
            for (uint32_t k = 0; k < a; ++k) sum -= M[b * (b + 1) / 2 + k] * M[a * (a + 1) / 2 + k];
            if (a == b)
                M[b * (b + 1) / 2 + b] = sum > 0.0 ? sqrt(sum) : 0.0;
            else
                M[b * (b + 1) / 2 + a] = M[a * (a + 1) / 2 + a] > 0.0 ? sum / M[a * (a + 1) / 2 + a] : 0.0;
        }
    }
    // forward C y = e_0, backward C^T g = y
    double y[kFsaiRegs], g[kFsaiRegs];
#pragma unroll
    for (uint32_t b = 0; b < kFsaiRegs; ++b) {
        double sum = b == 0 ? 1.0 : 0.0;
#pragma unroll
        for (uint32_t k = 0; k < b; ++k) sum -= M[b * (b + 1) / 2 + k] * y[k];
        const double d = M[b * (b + 1) / 2 + b];
        y[b] = d > 0.0 ? sum / d : 0.0;
    }
#pragma unroll
    for (int b = (int)kFsaiRegs - 1; b >= 0; --b) {
        double sum = y[b];
#pragma unroll
        for (uint32_t k = (uint32_t)b + 1; k < kFsaiRegs; ++k) sum -= M[k * (k + 1) / 2 + b] * g[k];
        const double d = M[b * (b + 1) / 2 + b];
        g[b] = d > 0.0 ? sum / d : 0.0;
    }
    if (!(g[0] > 0.0)) {
        // not positive (singular local block): Jacobi row
        const double a00 = M[0] > 0.0 ? M[0] * M[0] : 1.0;
        g[0] = 1.0 / a00;
#pragma unroll
        for (uint32_t k = 1; k < kFsaiRegs; ++k) g[k] = 0.0;
    }
    const double sc = 1.0 / sqrt(g[0]);
    diag[i] = (V)(g[0] * sc);
    // L_ji (j > i) into the up slot of (i, j) and the lo slot of (j, i)
    uint32_t k = 1;
    for (uint32_t e = row[i]; e < row[i + 1]; ++e)
        if (col[e] > i) {
            const V v = (V)(k < m ? g[k] * sc : 0.0);
            up_val[slot[e]] = v;
            const uint32_t f = csr_find(row, col, col[e], i);
            if (f != ~0u) lo_val[slot[f]] = v;
            ++k;
        }
}

}  // namespace

template <class V> void factor_structure(const DevCsr<double>& A, LowerFactor<V>& L, hipStream_t s) {
    L.n = A.n;
    L.slot.resize(A.nnz ? A.nnz : 1);
    build_sell<V>(A.n, A.row.p, A.col.p, A.val.p, nullptr, kSellStandard, kSellUpper, L.up, s, L.slot.p);
    build_sell<V>(A.n, A.row.p, A.col.p, A.val.p, nullptr, kSellStandard, kSellLower, L.lo, s, L.slot.p);
    L.diag.resize(A.n ? A.n : 1);
}

template <class V> void fsai_factor(const DevCsr<double>& A, LowerFactor<V>& L, hipStream_t s) {
    factor_structure(A, L, s);
    if (!A.n) return;
    hipLaunchKernelGGL((fsai_kernel<V>), dim3(grid_for(A.n, 256)), dim3(256), 0, s, A.n, A.row.p, A.col.p, A.val.p,
                       L.up.len.p, L.slot.p, L.up.val.p, L.lo.val.p, L.diag.p);
}

template void factor_structure<float>(const DevCsr<double>&, LowerFactor<float>&, hipStream_t);
template void factor_structure<double>(const DevCsr<double>&, LowerFactor<double>&, hipStream_t);
template void fsai_factor<float>(const DevCsr<double>&, LowerFactor<float>&, hipStream_t);
template void fsai_factor<double>(const DevCsr<double>&, LowerFactor<double>&, hipStream_t);

}  // namespace stream_amd
