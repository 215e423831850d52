// Factored approximate-inverse preconditioners M^-1 = L L^T with L lower
// triangular on the pattern of A (no counterpart in the reference, whose
// only preconditioner is Jacobi: include/numerics/preconditioner.hpp:60-72).
//
// Two producers of L:
//   * classical FSAI (Kolotilina-Yeremin) on the upper pattern: for row i,
//     P = {i} U {j > i : A_ij in pattern}; solve A[P,P] g = e_i exactly in
//     FP64 (dense Cholesky, |P| <= kFsaiMax) and scale g by 1/sqrt(g_i).
//     Then G (upper, rows g) satisfies diag(G A G^T) = 1 and L = G^T.
//   * the learned message-passing network (mpnn.hip), which writes the same
//     slots.
// Storage (LowerFactor): up = strict upper entries of row i (L_ji, j > i),
// lo = strict lower entries of row i (L_ij, j < i), diag = L_ii; both as
// packed SELL-64 copies of A's pattern, so the two half products of the
// apply are coalesced like the SpMV.  Producers write each value L_ji
// through LowerFactor::slot (CSR entry -> slot of its half) into the up slot
// of (i, j) and the lo slot of (j, i).
#pragma once
#include "numerics.hpp"

namespace stream_amd {

constexpr uint32_t kFsaiMax = 16;

// Build the lo/up SELL structures of L on A's pattern (values copied from A
// until a producer overwrites them) and the CSR entry -> slot map.
template <class V> void factor_structure(const DevCsr<double>& A, LowerFactor<V>& L, hipStream_t s);

// Fill L with the classical FSAI factor of A (FP64 solves, stored in V).
template <class V> void fsai_factor(const DevCsr<double>& A, LowerFactor<V>& L, hipStream_t s);

// Position of entry (j, i) in row j of a CSR whose rows have ascending
// columns, or ~0u when the pattern has no such entry.  A producer writing
// L_ji into the up slot of (i, j) writes the lo slot of (j, i) with it.
__device__ __forceinline__ uint32_t csr_find(const uint32_t* __restrict__ row, const uint32_t* __restrict__ col,
                                             uint32_t j, uint32_t i) {
    uint32_t a = row[j], b = row[j + 1];
    const uint32_t end = b;
    while (a < b) {
        const uint32_t m = (a + b) >> 1;
        if (col[m] < i)
            a = m + 1;
        else
            b = m;
    }
    return (a < end && col[a] == i) ? a : ~0u;
}

}  // namespace stream_amd
