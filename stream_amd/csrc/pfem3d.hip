// P1 pressure Laplacian on the 3D alpha-shape tetrahedra and the 3D
// mesh-and-solve step (no reference counterpart: the reference has no FEM
// module and never connects mesh and solve).
//
// Row i is assembled by node i's thread from the tetrahedra incident to i,
// visited in increasing element index, so the contributions to entry (i, j)
// are summed in the same order by rows i and j and the matrix is bitwise
// symmetric.  Per tetrahedron, with its vertices in ascending id p0..p3:
//   e_k = p_k - p0, g1 = e2 x e3, g2 = e3 x e1, g3 = e1 x e2,
//   g0 = -(g1 + g2 + g3), det = e1 . g1,
//   K_ab = (g_a . g_b) / (6 |det|)   (= V grad phi_a . grad phi_b),
//   lumped mass |det| / 24 per vertex.
// Diagonal = -(sum of the row's off-diagonals in ascending column order);
// Dirichlet p = 0 on free-surface nodes (identity rows, zeroed columns);
// b_i = f_i * mass_i.  The incident-element lists are a counting sort of
// the element array (per-node lists sorted after an atomic fill).
#include <cmath>
#include <cstring>
#include <vector>

#include "ashape3d.hpp"
#include "mpnn.hpp"
#include "numerics.hpp"
#include "precond.hpp"

namespace stream_amd {

namespace {

// Node -> tetrahedron incidence, counted and filled block by block: a
// block's 256 tetrahedra (at most 1024 distinct vertices) are first gathered
// in an LDS table of kIncSlots slots by open addressing (more slots than
// keys, so a probe ends at the vertex's slot or at an empty one; vertex ids
// are below n, never the empty marker), then each distinct vertex takes one
// global atomic for the block: consecutive tetrahedra share most vertices.
// The order of a node's list depends on the atomics' order; the column-set
// kernel sorts every list by tetrahedron id.
constexpr uint32_t kIncBlock = 256;
constexpr uint32_t kIncSlots = 2048;
static_assert(kIncSlots > 4 * kIncBlock, "the table must hold every vertex of a block with a slot to spare");

__device__ __forceinline__ uint32_t inc_slot(uint32_t v, uint32_t* hkey) {
    uint32_t h = (v * 2654435761u) >> 21;  // 11 bits: kIncSlots = 2048
    for (;;) {
        const uint32_t old = atomicCAS(&hkey[h], ~0u, v);
        if (old == ~0u || old == v) return h;
        h = (h + 1u) & (kIncSlots - 1u);
    }
}

__device__ __forceinline__ void inc_table_clear(uint32_t* hkey, uint32_t* hcnt) {
    for (uint32_t j = threadIdx.x; j < kIncSlots; j += kIncBlock) {
        hkey[j] = ~0u;
        hcnt[j] = 0u;
    }
}

__global__ __launch_bounds__(kIncBlock) void tet_count_kernel(uint32_t ne, const uint32_t* __restrict__ elem,
                                                              uint32_t* __restrict__ cnt) {
    __shared__ uint32_t hkey[kIncSlots], hcnt[kIncSlots];
    inc_table_clear(hkey, hcnt);
    __syncthreads();
    const uint32_t t = blockIdx.x * kIncBlock + threadIdx.x;
    if (t < ne) {
        const uint4 e = reinterpret_cast<const uint4*>(elem)[t];
        const uint32_t v[4] = {e.x, e.y, e.z, e.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) atomicAdd(&hcnt[inc_slot(v[k], hkey)], 1u);
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < kIncSlots; j += kIncBlock)
        if (hkey[j] != ~0u) atomicAdd(&cnt[hkey[j]], hcnt[j]);
}

// Each distinct vertex of the block reserves its block count of places in
// its list with one atomic; a reference takes the reserved run's start +
// its rank among the block's references to that vertex.
__global__ __launch_bounds__(kIncBlock) void tet_fill_kernel(uint32_t ne, const uint32_t* __restrict__ elem,
                                                             const uint32_t* __restrict__ toff,
                                                             uint32_t* __restrict__ cur, uint32_t* __restrict__ inc) {
    __shared__ uint32_t hkey[kIncSlots], hcnt[kIncSlots];
    inc_table_clear(hkey, hcnt);
    __syncthreads();
    const uint32_t t = blockIdx.x * kIncBlock + threadIdx.x;
    const bool live = t < ne;
    uint32_t h[4] = {0u, 0u, 0u, 0u}, r[4] = {0u, 0u, 0u, 0u};
    if (live) {
        const uint4 e = reinterpret_cast<const uint4*>(elem)[t];
        const uint32_t v[4] = {e.x, e.y, e.z, e.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            h[k] = inc_slot(v[k], hkey);
            r[k] = atomicAdd(&hcnt[h[k]], 1u);
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < kIncSlots; j += kIncBlock) {
        const uint32_t v = hkey[j];
        if (v != ~0u) hcnt[j] = toff[v] + atomicAdd(&cur[v], hcnt[j]);  // the run's start replaces the count
    }
    __syncthreads();
    if (live) {
#pragma unroll
        for (int k = 0; k < 4; ++k) inc[hcnt[h[k]] + r[k]] = t;
    }
}

// x on lane l, where l is the same on every lane of the wavefront: a scalar
// read of one lane's register (v_readlane) instead of a cross-lane permute
// through the LDS unit; the value is the one __shfl(x, l) returns.
__device__ __forceinline__ uint32_t lane_u32(uint32_t x, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, (int)l);
}
__device__ __forceinline__ double lane_f64(double x, uint32_t l) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const uint64_t lo = lane_u32((uint32_t)b, l), hi = lane_u32((uint32_t)(b >> 32), l);
    return __longlong_as_double((long long)(lo | (hi << 32)));
}

__device__ __forceinline__ void sort4(uint32_t* v) {
    uint32_t t;
    if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
    if (v[2] > v[3]) { t = v[2]; v[2] = v[3]; v[3] = t; }
    if (v[0] > v[2]) { t = v[0]; v[0] = v[2]; v[2] = t; }
    if (v[1] > v[3]) { t = v[1]; v[1] = v[3]; v[3] = t; }
    if (v[1] > v[2]) { t = v[1]; v[1] = v[2]; v[2] = t; }
}

// Sort node i's incident elements and build its sorted unique column list
// (i and the vertices of its elements) in tmp[3*toff[i] + i ...].
// Column set of row i (serial form, one thread): the incident tetrahedra in
// ascending id order, then every vertex of them inserted into a sorted,
// duplicate-free list that starts with i.
__device__ void tet_cols_serial(uint32_t i, const uint32_t* __restrict__ elem, const uint32_t* __restrict__ toff,
                                uint32_t* __restrict__ inc, uint32_t* __restrict__ tmp, uint32_t* __restrict__ ncol) {
    const uint32_t b0 = toff[i], b1 = toff[i + 1];
    for (uint32_t k = b0 + 1; k < b1; ++k) {
        const uint32_t x = inc[k];
        uint32_t p = k;
        while (p > b0 && inc[p - 1] > x) {
            inc[p] = inc[p - 1];
            --p;
        }
        inc[p] = x;
    }
    uint32_t* c = tmp + 3ull * b0 + i;
    uint32_t cnt = 1;
    c[0] = i;
    for (uint32_t k = b0; k < b1; ++k) {
        const uint32_t t = inc[k];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t v = elem[4ull * t + q];
            // sorted insert if absent
            uint32_t lo = 0, hi = cnt;
            while (lo < hi) {
                const uint32_t m = (lo + hi) >> 1;
                if (c[m] < v)
                    lo = m + 1;
                else
                    hi = m;
            }
            if (lo < cnt && c[lo] == v) continue;
            for (uint32_t p = cnt; p > lo; --p) c[p] = c[p - 1];
            c[lo] = v;
            ++cnt;
        }
    }
    ncol[i] = cnt;
}

// The same column sets, one wavefront per row: the (at most 64) incident
// tetrahedra sit one per lane and are placed by their rank among the
// row's ids; the 3 m vertices other than i and i itself (at most 193 keys,
// 4 slots per lane) are reduced to their distinct values by two passes over
// all keys broadcast lane by lane: a key is kept if no earlier key is equal, and a
// kept key goes to the position given by the number of kept keys below
// it.  The output (sorted tetrahedra, sorted distinct columns, count)
// equals the serial form's; rows with more than 64 incident tetrahedra
// take the serial form on one lane.
// Key dedupe of one row with NS 64-key slots per lane (3 m + 1 <= 64 NS).
template <uint32_t NS>
__device__ __forceinline__ void tet_cols_row(uint32_t i, uint32_t b0, uint32_t m, uint32_t t, uint32_t rt,
                                             uint32_t lane, const uint32_t* __restrict__ elem,
                                             uint32_t* __restrict__ inc, uint32_t* __restrict__ tmp,
                                             uint32_t* __restrict__ ncol) {
    // keys: position p = slot * 64 + lane; p < 3 m: the (p % 3)-th vertex
    // other than i of the tetrahedron on lane p / 3; p = 3 m: i itself
    const uint32_t nk = 3 * m + 1;
    uint32_t key[NS];
    bool first[NS];
#pragma unroll
    for (uint32_t sl = 0; sl < NS; ++sl) {
        const uint32_t p = sl * 64 + lane;
        const uint32_t j = p / 3, q = p - 3 * j;
        const uint32_t tj = (uint32_t)__shfl((int)t, (int)(j & 63u), 64);
        uint32_t kv = ~0u;
        if (p < 3 * m) {
            uint32_t c = 0;
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) {
                const uint32_t v = elem[4ull * tj + u];
                if (v != i) {
                    if (c == q) kv = v;
                    ++c;
                }
            }
        } else if (p == 3 * m) {
            kv = i;
        }
        key[sl] = kv;
        first[sl] = p < nk;
    }
    if (lane < m) inc[b0 + rt] = t;  // every lane has read its tetrahedron
    // pass 1: a key is kept when no earlier position holds the same value
#pragma unroll
    for (uint32_t S = 0; S < NS; ++S) {
        if (S * 64 >= nk) break;
        const uint32_t lim = nk - S * 64 < 64 ? nk - S * 64 : 64;
        for (uint32_t L = 0; L < lim; ++L) {
            const uint32_t kP = lane_u32(key[S], L);
            const uint32_t P = S * 64 + L;
#pragma unroll
            for (uint32_t sl = 0; sl < NS; ++sl)
                if (kP == key[sl] && P < sl * 64 + lane) first[sl] = false;
        }
    }
    // pass 2: a kept key's position = number of kept keys below it
    uint32_t rank[NS];
#pragma unroll
    for (uint32_t sl = 0; sl < NS; ++sl) rank[sl] = 0;
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t S = 0; S < NS; ++S) {
        if (S * 64 >= nk) break;
        cnt += (uint32_t)__popcll(__ballot(first[S]));
        const uint32_t lim = nk - S * 64 < 64 ? nk - S * 64 : 64;
        for (uint32_t L = 0; L < lim; ++L) {
            const uint32_t kP = lane_u32(key[S], L);
            const bool fP = lane_u32(first[S] ? 1u : 0u, L) != 0u;
#pragma unroll
            for (uint32_t sl = 0; sl < NS; ++sl) rank[sl] += (fP && kP < key[sl]) ? 1u : 0u;
        }
    }
    uint32_t* c = tmp + 3ull * b0 + i;
#pragma unroll
    for (uint32_t sl = 0; sl < NS; ++sl)
        if (first[sl]) c[rank[sl]] = key[sl];
    if (lane == 0) ncol[i] = cnt;
}

__global__ __launch_bounds__(256) void tet_cols_kernel(uint32_t n, const uint32_t* __restrict__ elem,
                                                       const uint32_t* __restrict__ toff, uint32_t* __restrict__ inc,
                                                       uint32_t* __restrict__ tmp, uint32_t* __restrict__ ncol) {
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t i0 = blockIdx.x * 4 + (threadIdx.x >> 6); i0 < n; i0 += gridDim.x * 4) {
        const uint32_t i = __builtin_amdgcn_readfirstlane(i0);
        const uint32_t b0 = toff[i], b1 = toff[i + 1];
        const uint32_t m = b1 - b0;
        if (m > 64) {
            if (lane == 0) tet_cols_serial(i, elem, toff, inc, tmp, ncol);
            continue;
        }
        const uint32_t t = lane < m ? inc[b0 + lane] : ~0u;
        // rank of this lane's tetrahedron among the row's (ids are distinct)
        uint32_t rt = 0;
        for (uint32_t l = 0; l < m; ++l) rt += lane_u32(t, l) < t ? 1u : 0u;
        // only the slots that hold keys are walked
        const uint32_t ns = (3 * m + 1 + 63) / 64;
        if (ns <= 1)
            tet_cols_row<1>(i, b0, m, t, rt, lane, elem, inc, tmp, ncol);
        else if (ns == 2)
            tet_cols_row<2>(i, b0, m, t, rt, lane, elem, inc, tmp, ncol);
        else if (ns == 3)
            tet_cols_row<3>(i, b0, m, t, rt, lane, elem, inc, tmp, ncol);
        else
            tet_cols_row<4>(i, b0, m, t, rt, lane, elem, inc, tmp, ncol);
    }
}

struct P1TetArgs {
    uint32_t n;
    const Sph4<double>* P;
    const uint32_t* elem;
    const uint32_t* toff;
    const uint32_t* inc;
    const uint32_t* tmp;
    const uint8_t* bnd;
    const double* f;
    uint32_t* arow;  // filled by the scan
    uint32_t* acol;
    double* aval;
    double* idiag;
    float* idiag32;
    double* b;
};

// Terms of tetrahedron t at row i: the gradients of the P1 basis (cross
// products of the edges from the lowest vertex), the stiffness
// K_i,other = g_i . g_other / (6 |det|) of the three edges through i, in the
// order of the vertex pairs (0,1) (0,2) (0,3) (1,2) (1,3) (2,3), and the
// lumped mass share (|det| / 6) / 4.
__device__ __forceinline__ void tet_terms(const P1TetArgs& a, uint32_t i, uint32_t t, uint32_t (&oth)[3],
                                          double (&kv)[3], double& mass_term) {
    uint32_t v[4] = {a.elem[4ull * t], a.elem[4ull * t + 1], a.elem[4ull * t + 2], a.elem[4ull * t + 3]};
    sort4(v);
    double p[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const Sph4<double> s = a.P[v[q]];
        p[q][0] = s.c[0];
        p[q][1] = s.c[1];
        p[q][2] = s.c[2];
    }
    double e[3][3];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) e[q][c] = p[q + 1][c] - p[0][c];
    double g[4][3];
    auto cross = [](const double* x, const double* y, double* o) {
        o[0] = x[1] * y[2] - x[2] * y[1];
        o[1] = x[2] * y[0] - x[0] * y[2];
        o[2] = x[0] * y[1] - x[1] * y[0];
    };
    cross(e[1], e[2], g[1]);
    cross(e[2], e[0], g[2]);
    cross(e[0], e[1], g[3]);
#pragma unroll
    for (int c = 0; c < 3; ++c) g[0][c] = -((g[1][c] + g[2][c]) + g[3][c]);
    const double det = e[0][0] * g[1][0] + e[0][1] * g[1][1] + e[0][2] * g[1][2];
    const double ad = fabs(det);
    const double den = 6.0 * ad;
    int li = 0;
#pragma unroll
    for (int q = 1; q < 4; ++q)
        if (v[q] == i) li = q;
    constexpr int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {1, 2, 3, 2, 3, 3};
    int o = 0;
#pragma unroll
    for (int kk = 0; kk < 6; ++kk) {
        if (pa[kk] != li && pb[kk] != li) continue;
        const double* x = g[pa[kk]];
        const double* y = g[pb[kk]];
        const double kvv = (x[0] * y[0] + x[1] * y[1] + x[2] * y[2]) / den;
        const uint32_t other = v[pa[kk] == li ? pb[kk] : pa[kk]];
#pragma unroll
        for (int q = 0; q < 3; ++q)
            if (q == o) {
                oth[q] = other;
                kv[q] = kvv;
            }
        ++o;
    }
    mass_term = (ad / 6.0) / 4.0;
}

// Row i by one thread (rows with more than 64 tetrahedra or columns): the
// terms of the incident tetrahedra in incidence order, each added to its
// column (binary search in the sorted row), then the Dirichlet rows and
// columns, the diagonal as minus the off-diagonal sum, and the load.
__device__ void p1_tet_row_serial(const P1TetArgs& a, uint32_t i) {
    const uint32_t r0 = a.arow[i], cnt = a.arow[i + 1] - r0;
    const uint32_t* src = a.tmp + 3ull * a.toff[i] + i;
    uint32_t* col = a.acol + r0;
    double* val = a.aval + r0;
    for (uint32_t k = 0; k < cnt; ++k) {
        col[k] = src[k];
        val[k] = 0.0;
    }
    auto find = [&](uint32_t v) -> uint32_t {
        uint32_t lo = 0, hi = cnt;
        while (hi - lo > 1) {
            const uint32_t m = (lo + hi) >> 1;
            if (col[m] <= v)
                lo = m;
            else
                hi = m;
        }
        return lo;
    };
    double mass = 0.0;
    for (uint32_t k = a.toff[i]; k < a.toff[i + 1]; ++k) {
        uint32_t oth[3];
        double kv[3], mt;
        tet_terms(a, i, a.inc[k], oth, kv, mt);
#pragma unroll
        for (int q = 0; q < 3; ++q) val[find(oth[q])] += kv[q];
        mass += mt;
    }
    const bool bi = a.bnd[i] != 0;
    double diag = 0.0;
    uint32_t dpos = 0;
    for (uint32_t k = 0; k < cnt; ++k) {
        if (col[k] == i) {
            dpos = k;
            continue;
        }
        diag -= val[k];
    }
    for (uint32_t k = 0; k < cnt; ++k) {
        double vv;
        if (k == dpos)
            vv = bi ? 1.0 : diag;
        else
            vv = (bi || a.bnd[col[k]]) ? 0.0 : val[k];
        val[k] = vv;
    }
    const double dg = bi ? 1.0 : diag;
    const double id = dg != 0.0 ? 1.0 / dg : 1.0;
    a.idiag[i] = id;
    a.idiag32[i] = (float)id;
    a.b[i] = bi ? 0.0 : a.f[i] * mass;
}

// One wavefront per row: lane t computes the terms of the row's t-th
// incident tetrahedron, lane c holds column c of the sorted row and adds the
// terms that fall on it in tetrahedron order (each tetrahedron reaches a
// column at most once), so every entry is the same sum in the same order as
// the serial row; the mass and the diagonal are summed in the serial order
// too.  Larger rows take the serial code on lane 0.
__global__ __launch_bounds__(256) void p1_tet_kernel(P1TetArgs a) {
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t i0 = blockIdx.x * 4 + (threadIdx.x >> 6); i0 < a.n; i0 += gridDim.x * 4) {
        const uint32_t i = __builtin_amdgcn_readfirstlane(i0);
        const uint32_t b0 = a.toff[i], m = a.toff[i + 1] - b0;
        const uint32_t r0 = a.arow[i], cnt = a.arow[i + 1] - r0;
        if (m > 64 || cnt > 64) {
            if (lane == 0) p1_tet_row_serial(a, i);
            continue;
        }
        uint32_t oth[3] = {~0u, ~0u, ~0u};
        double kv[3] = {0.0, 0.0, 0.0}, mt = 0.0;
        if (lane < m) tet_terms(a, i, a.inc[b0 + lane], oth, kv, mt);
        const uint32_t cc = lane < cnt ? a.tmp[3ull * b0 + i + lane] : ~0u;
        double acc = 0.0, mass = 0.0;
        for (uint32_t t = 0; t < m; ++t) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const uint32_t o = lane_u32(oth[q], t);
                const double k = lane_f64(kv[q], t);
                if (o == cc) acc += k;
            }
            mass += lane_f64(mt, t);
        }
        const bool bi = a.bnd[i] != 0;
        double diag = 0.0;
        uint32_t dpos = 0;
        for (uint32_t k = 0; k < cnt; ++k) {
            const uint32_t ck = lane_u32(cc, k);
            const double vk = lane_f64(acc, k);
            if (ck == i)
                dpos = k;
            else
                diag -= vk;
        }
        if (lane < cnt) {
            double vv;
            if (lane == dpos)
                vv = bi ? 1.0 : diag;
            else
                vv = (bi || a.bnd[cc]) ? 0.0 : acc;
            a.acol[r0 + lane] = cc;
            a.aval[r0 + lane] = vv;
        }
        if (lane == 0) {
            const double dg = bi ? 1.0 : diag;
            const double id = dg != 0.0 ? 1.0 / dg : 1.0;
            a.idiag[i] = id;
            a.idiag32[i] = (float)id;
            a.b[i] = bi ? 0.0 : a.f[i] * mass;
        }
    }
}

template <class T>
__global__ __launch_bounds__(256) void permute3_fwd_kernel(const T* __restrict__ in, const uint32_t* __restrict__ perm,
                                                           uint32_t n, T* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[perm[i]];
}
template <class T>
__global__ __launch_bounds__(256) void permute3_bwd_kernel(const T* __restrict__ in, const uint32_t* __restrict__ perm,
                                                           uint32_t n, T* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[perm[i]] = in[i];
}

}  // namespace

namespace {
// Divergence-free 3D velocity u(x) = sum_m cos(k_m . x + phi_m) b_m with
// b_m = k_m x a_m, the curl of the vector potential sum_m a_m sin(k_m . x +
// phi_m).  Rows (kx, ky, kz, phi, bx, by, bz).
struct Modes3 {
    uint32_t n = 0;
    double m[32][7];
};

// One forward-Euler step x += dt u(x) on the input-order points.
__global__ __launch_bounds__(256) void advect3_kernel(uint32_t n, Modes3 md, double dt,
                                                      Sph4<double>* __restrict__ pts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Sph4<double> p = pts[i];
    double u[3] = {0.0, 0.0, 0.0};
    for (uint32_t k = 0; k < md.n; ++k) {
        const double* r = md.m[k];
        const double c = cos(r[0] * p.c[0] + r[1] * p.c[1] + r[2] * p.c[2] + r[3]);
        u[0] += c * r[4];
        u[1] += c * r[5];
        u[2] += c * r[6];
    }
    p.c[0] += dt * u[0];
    p.c[1] += dt * u[1];
    p.c[2] += dt * u[2];
    pts[i] = p;
}
}  // namespace

struct Step3D {
    Ash3D<double> mesh;
    DBuf<Sph4<double>> pts;  // input order
    DBuf<double> f_in, f_m, b, x, p_out;
    DevCsr<double> A;
    DBuf<uint32_t> tcnt, toff, tcur, inc, tmp, ncol;
    DBuf<uint64_t> scan_ws;
    Sell<double> S64;
    Sell<float> S32;
    LowerFactor<float> L32;
    DBuf<float> idiag32;
    DBuf<double> idiag;
    CgWork work;
    SolveStats last;
    double t_mesh = 0, t_asm = 0, t_prep = 0, t_solve = 0;
    uint32_t n = 0;
    PrecLearned* learned = nullptr;  // network of solver 3 (not owned)
    bool learned_fallback = false;   // the last learned factor was rejected

    void advect(const Modes3& md, double dt, uint32_t steps) {
        hipStream_t s = ctx().stream;
        for (uint32_t k = 0; k < steps; ++k)
            hipLaunchKernelGGL(advect3_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, md, dt, pts.p);
        SA_HIP(hipStreamSynchronize(s));
    }

    void set_points(uint32_t n_, const double* h_xyz, const double* h_alpha, const double* h_f) {
        n = n_;
        std::vector<Sph4<double>> h(n);
        for (uint32_t i = 0; i < n; ++i) {
            h[i].c[0] = h_xyz[3ull * i];
            h[i].c[1] = h_xyz[3ull * i + 1];
            h[i].c[2] = h_xyz[3ull * i + 2];
            h[i].r = h_alpha[i];
        }
        pts.resize(n ? n : 1);
        h2d(pts.p, h.data(), n);
        f_in.resize(n ? n : 1);
        h2d(f_in.p, h_f, n);
    }

    void assemble(hipStream_t s) {
        const uint32_t ne = mesh.n_elems();
        A.n = n;
        tcnt.resize(n + 1);
        toff.resize(n + 1);
        tcur.resize(n + 1);
        ncol.resize(n + 1);
        A.row.resize(n + 1);
        scan_ws.resize(scan_workspace_words(n + 1));
        fill<uint32_t>(tcnt.p, n + 1, 0u, s);
        fill<uint32_t>(tcur.p, n + 1, 0u, s);
        if (ne)
            hipLaunchKernelGGL(tet_count_kernel, dim3((ne + kIncBlock - 1) / kIncBlock), dim3(kIncBlock), 0, s, ne,
                               mesh.elemglob.p, tcnt.p);
        exclusive_scan_u32(tcnt.p, toff.p, n, scan_ws.p, s);
        inc.resize(4ull * ne + 1);
        tmp.resize(12ull * ne + n + 1);
        if (ne)
            hipLaunchKernelGGL(tet_fill_kernel, dim3((ne + kIncBlock - 1) / kIncBlock), dim3(kIncBlock), 0, s, ne,
                               mesh.elemglob.p, toff.p, tcur.p, inc.p);
        hipLaunchKernelGGL(tet_cols_kernel, dim3(grid_for(n, 4)), dim3(256), 0, s, n, mesh.elemglob.p, toff.p,
                           inc.p, tmp.p, ncol.p);
        exclusive_scan_u32(ncol.p, A.row.p, n, scan_ws.p, s);
        A.nnz = read_scalar(A.row.p + n);
        A.col.resize(A.nnz ? A.nnz : 1);
        A.val.resize(A.nnz ? A.nnz : 1);
        idiag.resize(n);
        idiag32.resize(n);
        b.resize(n);
        f_m.resize(n);
        x.resize(n);
        hipLaunchKernelGGL((permute3_fwd_kernel<double>), dim3(grid_for(n, 256)), dim3(256), 0, s, f_in.p,
                           mesh.lbvh.perm, n, f_m.p);
        P1TetArgs a;
        a.n = n;
        a.P = mesh.lbvh.obj_m.p;
        a.elem = mesh.elemglob.p;
        a.toff = toff.p;
        a.inc = inc.p;
        a.tmp = tmp.p;
        a.bnd = mesh.node_is_bnd.p;
        a.f = f_m.p;
        a.arow = A.row.p;
        a.acol = A.col.p;
        a.aval = A.val.p;
        a.idiag = idiag.p;
        a.idiag32 = idiag32.p;
        a.b = b.p;
        hipLaunchKernelGGL(p1_tet_kernel, dim3(grid_for(n, 4)), dim3(256), 0, s, a);
    }

    // Point the network's graph at this step's system (after assembly).
    void bind_graph(MpnnGraph& g, hipStream_t s, bool transpose = true) {
        g.n = n;
        g.nnz = A.nnz;
        g.row = A.row.p;
        g.col = A.col.p;
        g.val = A.val.p;
        g.xy = reinterpret_cast<const double*>(mesh.lbvh.obj_m.p);  // Sph4: c[0..2] first
        g.xy_stride = sizeof(Sph4<double>) / sizeof(double);
        g.dim = 3;
        g.bnd = mesh.node_is_bnd.p;
        mpnn_prepare(g, s, transpose);
    }

    // SELL copy and factor for the solver; the learned factor is rejected
    // (Jacobi used instead) when a value is non-finite or a diagonal is not
    // positive.
    void prepare(int solver, hipStream_t s) {
        learned_fallback = false;
        if (solver == 0) {
            build_sell<double>(n, A.row.p, A.col.p, A.val.p, nullptr, kSellStandard, kSellOffDiag, S64, s);
        } else {
            build_sell<float>(n, A.row.p, A.col.p, A.val.p, nullptr, kSellRowSum, kSellOffDiag, S32, s);
            if (solver == 2) fsai_factor<float>(A, L32, s);
            if (solver == 3) {
                if (!learned) throw std::invalid_argument("solver 3 needs a network (PfemStep3D_set_learned)");
                factor_structure<float>(A, L32, s);
                bind_graph(learned->graph, s, false);  // inference: no transpose map
                learned->net.forward(learned->graph, s, false);
                learned->net.pack(learned->graph, L32, s);
                const uint32_t bad = read_scalar(ctx().d_flags + kFlagLearned);
                if (bad) {
                    fill<uint32_t>(ctx().d_flags + kFlagLearned, 1, 0u, s);
                    learned_fallback = true;
                }
            }
        }
    }

    void mesh_and_assemble(int bits, int solver) {
        hipStream_t s = ctx().stream;
        hipEvent_t ev[4];
        for (auto& e : ev) SA_HIP(hipEventCreate(&e));
        SA_HIP(hipEventRecord(ev[0], s));
        mesh.bits = bits;
        mesh.set_nodes_device(n, pts.p);
        mesh.init(s);
        mesh.compute(s);
        SA_HIP(hipEventRecord(ev[1], s));
        assemble(s);
        SA_HIP(hipEventRecord(ev[2], s));
        if (solver >= 0) prepare(solver, s);
        SA_HIP(hipEventRecord(ev[3], s));
        SA_HIP(hipEventSynchronize(ev[3]));
        float t0 = 0, t1 = 0, t2 = 0;
        SA_HIP(hipEventElapsedTime(&t0, ev[0], ev[1]));
        SA_HIP(hipEventElapsedTime(&t1, ev[1], ev[2]));
        SA_HIP(hipEventElapsedTime(&t2, ev[2], ev[3]));
        t_mesh = t0;
        t_asm = t1;
        t_prep = t2;
        for (auto& e : ev) SA_HIP(hipEventDestroy(e));
    }

    void solve(int solver, double rtol, uint32_t maxit) {
        if (solver == 3 && learned_fallback) solver = 1;
        hipStream_t s = ctx().stream;
        fill<double>(x.p, n, 0.0, s);
        CgOptions o;
        o.rtol = rtol;
        o.maxit = maxit;
        o.batch = tuning().batch;
        if (solver == 0) {
            cg_solve<double>(A, S64, false, kPrecJacobi, idiag.p, nullptr, b.p, x.p, o, work, last, s);
        } else {
            o.replace_factor = tuning().replace_factor;
            if (solver == 1)
                cg_solve<float>(A, S32, true, kPrecJacobi, idiag32.p, nullptr, b.p, x.p, o, work, last, s);
            else
                cg_solve<float>(A, S32, true, kPrecLLt, nullptr, &L32, b.p, x.p, o, work, last, s);
        }
        t_solve = last.ms;
    }
};

}  // namespace stream_amd

using namespace stream_amd;

// ===========================================================================
// PfemStep3D C-ABI (stream_numerics.h).
// ===========================================================================
struct PfemStep3D {
    Step3D s;
    int bits = 21;
};

namespace {
template <class F> int32_t g3(F&& f) { return guarded(std::forward<F>(f)); }
}  // namespace

extern "C" {

PfemStep3D* PfemStep3D_create(void) {
    try {
        (void)ctx();
        return new PfemStep3D();
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return nullptr;
    }
}
void PfemStep3D_destroy(PfemStep3D* h) { delete h; }

// Points (row-major xyz [3N]), alpha [N], nodal source f [N], input order.
int32_t PfemStep3D_set_points(PfemStep3D* h, uint32_t n, const double* xyz, const double* alpha, const double* f) {
    return g3([&] { h->s.set_points(n, xyz, alpha, f); });
}
int32_t PfemStep3D_set_morton_bits(PfemStep3D* h, int32_t bits) {
    return g3([&] {
        if (bits < 1 || bits > 21) throw std::invalid_argument("3D Morton bits per axis must be in [1, 21]");
        h->bits = bits;
    });
}

// One 3D step: LBVH + alpha-shape tetrahedra + P1 assembly + CG.  solver:
// 0 = FP64 Jacobi-PCG, 1 = mixed FP32/FP64 Jacobi-PCG, 2 = mixed PCG with the
// FSAI L L^T factor, 3 = mixed PCG with the learned L L^T factor (network
// set by PfemStep3D_set_learned).  stats [10] as PfemStep2D_run.
int32_t PfemStep3D_run(PfemStep3D* h, int32_t solver, double rtol, uint32_t maxit, double* stats) {
    return g3([&] {
        if (solver < 0 || solver > 3) throw std::invalid_argument("3D step solvers are 0, 1, 2 and 3");
        h->s.mesh_and_assemble(h->bits, solver);
        h->s.solve(solver, rtol, maxit);
        if (stats) {
            stats[0] = h->s.t_mesh;
            stats[1] = h->s.t_asm;
            stats[2] = h->s.t_solve;
            stats[3] = h->s.last.iters;
            stats[4] = h->s.last.true_rel;
            stats[5] = h->s.last.replacements;
            stats[6] = h->s.A.nnz;
            stats[7] = h->s.mesh.n_elems();
            stats[8] = h->s.t_prep;
            stats[9] = h->s.learned_fallback ? 1.0 : 0.0;
        }
        const uint32_t fl = read_scalar(ctx().d_flags + kFlagError);
        if (fl) {
            fill<uint32_t>(ctx().d_flags, 1, 0u, ctx().stream);
            throw std::runtime_error("device error flag raised during the 3D step");
        }
        if (h->s.last.true_rel > rtol) throw std::runtime_error("CG did not reach the requested residual");
    });
}

// Network used by solver 3 (not owned; nullptr clears it).
int32_t PfemStep3D_set_learned(PfemStep3D* h, PrecLearned* net) {
    return g3([&] { h->s.learned = net; });
}
// Point the network at the step's current (assembled) system.
int32_t PrecLearned_bind_step3d(PrecLearned* net, PfemStep3D* h) {
    return g3([&] {
        if (!net || !h) throw std::invalid_argument("PrecLearned_bind_step3d: null argument");
        h->s.bind_graph(net->graph, ctx().stream);
        mpnn_check_flags();
    });
}

// Advect the points `steps` forward-Euler steps of dt in the
// divergence-free field of n_modes rows (kx, ky, kz, phi, bx, by, bz):
// u(x) = sum cos(k . x + phi) b (b = k x a, the curl of a sin(k . x +
// phi)).  The source f is left as given.
int32_t PfemStep3D_advect(PfemStep3D* h, uint32_t n_modes, const double* modes, double dt, uint32_t steps) {
    return g3([&] {
        if (n_modes > 32) throw std::invalid_argument("at most 32 modes");
        Modes3 md;
        md.n = n_modes;
        for (uint32_t k = 0; k < n_modes; ++k)
            for (int c = 0; c < 7; ++c) md.m[k][c] = modes[7 * k + c];
        h->s.advect(md, dt, steps);
    });
}
// Current points (input order, row-major xyz [3N]).
int32_t PfemStep3D_get_points(PfemStep3D* h, double* xyz) {
    return g3([&] {
        std::vector<Sph4<double>> hp(h->s.n);
        d2h(hp.data(), h->s.pts.p, h->s.n);
        for (uint32_t i = 0; i < h->s.n; ++i)
            for (int c = 0; c < 3; ++c) xyz[3ull * i + c] = hp[i].c[c];
    });
}

// Mesh and assemble only; stats [2] (nullable): mesh ms, assembly ms.
int32_t PfemStep3D_assemble(PfemStep3D* h, double* stats) {
    return g3([&] {
        h->s.mesh_and_assemble(h->bits, -1);
        if (stats) {
            stats[0] = h->s.t_mesh;
            stats[1] = h->s.t_asm;
        }
        const uint32_t fl = read_scalar(ctx().d_flags + kFlagError);
        if (fl) {
            fill<uint32_t>(ctx().d_flags, 1, 0u, ctx().stream);
            throw std::runtime_error("device error flag raised during the 3D step");
        }
    });
}

int32_t PfemStep3D_get_solution(PfemStep3D* h, double* p) {
    return g3([&] {
        hipStream_t s = ctx().stream;
        const uint32_t n = h->s.n;
        h->s.p_out.resize(n ? n : 1);
        hipLaunchKernelGGL((permute3_bwd_kernel<double>), dim3(grid_for(n, 256)), dim3(256), 0, s, h->s.x.p,
                           h->s.mesh.lbvh.perm, n, h->s.p_out.p);
        d2h(p, h->s.p_out.p, n);
    });
}

uint32_t PfemStep3D_nnz(PfemStep3D* h) { return h ? h->s.A.nnz : 0u; }

// Assembled system in Morton order: row [N+1], col/val [nnz], b [N].
int32_t PfemStep3D_get_system(PfemStep3D* h, uint32_t* row, uint32_t* col, double* val, double* b) {
    return g3([&] {
        const uint32_t n = h->s.n, nnz = h->s.A.nnz;
        if (row) d2h(row, h->s.A.row.p, n + 1);
        if (col) d2h(col, h->s.A.col.p, nnz);
        if (val) d2h(val, h->s.A.val.p, nnz);
        if (b) d2h(b, h->s.b.p, n);
    });
}

// Mesh outputs (Morton ids): tetrahedra [4*nelem], node_is_bnd [N],
// Morton-ordered xyz [3N], permutation [N]; any may be null.
int32_t PfemStep3D_get_mesh(PfemStep3D* h, uint32_t* elem, uint8_t* node_is_bnd, double* xyz, uint32_t* perm) {
    return g3([&] {
        auto& m = h->s.mesh;
        const uint32_t n = m.n;
        if (elem) d2h(elem, m.elemglob.p, 4ull * m.n_elems());
        if (node_is_bnd) d2h(node_is_bnd, m.node_is_bnd.p, n);
        if (perm) d2h(perm, m.lbvh.perm, n);
        if (xyz) {
            std::vector<Sph4<double>> v(n);
            d2h(v.data(), m.lbvh.obj_m.p, n);
            for (uint32_t i = 0; i < n; ++i)
                for (int c = 0; c < 3; ++c) xyz[3 * i + c] = v[i].c[c];
        }
    });
}

uint32_t PfemStep3D_get_nelem(PfemStep3D* h) { return h ? h->s.mesh.n_elems() : 0u; }


// Nodal fields of the current points between input order and the Morton
// order of the step's mesh, matrix and solution (permuteForward /
// permuteBackward of the alpha-shape): forward out[i] = in[perm[i]],
// backward out[perm[i]] = in[i]; host arrays of n elements of elem_bytes.
// Needs a meshed step.
int32_t PfemStep3D_permute_forward(PfemStep3D* h, const void* in, void* out, uint32_t elem_bytes) {
    return g3([&] {
        if (!h || !h->s.mesh.inited || h->s.mesh.n != h->s.n) throw std::invalid_argument("PfemStep3D: no mesh for the current points");
        permute_host_elements(in, out, h->s.mesh.lbvh.perm, h->s.n, elem_bytes, true);
    });
}
int32_t PfemStep3D_permute_backward(PfemStep3D* h, const void* in, void* out, uint32_t elem_bytes) {
    return g3([&] {
        if (!h || !h->s.mesh.inited || h->s.mesh.n != h->s.n) throw std::invalid_argument("PfemStep3D: no mesh for the current points");
        permute_host_elements(in, out, h->s.mesh.lbvh.perm, h->s.n, elem_bytes, false);
    });
}

}  // extern "C"
