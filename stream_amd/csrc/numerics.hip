// Fused device-resident CG kernels and host drivers (see numerics.hpp).
#include <cmath>
#include <cstring>
#include <vector>

#include "numerics.hpp"

namespace stream_amd {

// Block size of the Jacobi-path CG kernels; the L L^T path's kernels are
// instantiated for 256, 512 and 1024 (Tuning::llt_threads).  Every kernel
// takes its block size NT as a template argument.
constexpr int kCgThreads = 256;
// Grid of every CG kernel: at most Tuning::cg_blocks blocks (<= 4096) with
// grid-stride rows; the last block folds at most 16 partials per thread,
// all loads in flight at once.
constexpr uint32_t kMaxCgBlocks = 4096;
constexpr uint32_t kMaxFoldGroups = kMaxCgBlocks / 8;  // smallest fold group: 8 blocks
constexpr uint32_t kCounterStride = 32;                // one 128-byte line per ticket counter
constexpr uint32_t kCounterWords = (1 + kMaxFoldGroups) * kCounterStride;
// Block partials handed from one kernel to the next (store_partials /
// sum_partials): the L L^T path's d.q at slot 0 and w.w, r'.r' at slots
// 1, 2 (stride kMaxCgBlocks); the Jacobi path's d.q after the grid folds'
// area, which the update kernel's fold must not overwrite while other
// blocks of it still read them.
constexpr size_t kJacobiDqSlot = 4ull * kMaxCgBlocks + 4ull * kMaxFoldGroups;

Tuning& tuning() {
    static Tuning t;
    return t;
}

void CgWork::resize(uint32_t n, uint32_t nthreads) {
    if (!(nthreads == 256 || nthreads == 512 || nthreads == 1024))
        throw std::invalid_argument("CG block size must be 256, 512 or 1024");
    threads = nthreads;
    blocks = (n + threads - 1) / threads;
    // cg_blocks caps the grid in 256-thread units: the same number of
    // threads whatever the block size
    uint32_t cap = tuning().cg_blocks < kMaxCgBlocks ? tuning().cg_blocks : kMaxCgBlocks;
    cap = cap * kCgThreads / threads;
    if (cap == 0) cap = 1;
    if (blocks > cap) blocks = cap;
    if (blocks == 0) blocks = 1;
    sc.resize(1);
    out2.resize(2);
    partials.resize(kJacobiDqSlot + kMaxCgBlocks);
    if (!counter.p) {
        counter.resize(kCounterWords);
        fill<uint32_t>(counter.p, kCounterWords, 0u, ctx().stream);
    }
}

// Block-group size of the two-level fold (0: one level), read by every
// fold; set through Stream_tune("fold_group", ...).
__constant__ uint32_t c_fold_group = 32;

void set_fold_group(uint32_t g) {
    if (!(g == 0 || g == 8 || g == 16 || g == 32 || g == 64))
        throw std::invalid_argument("fold_group must be 0, 8, 16, 32 or 64");
    (void)ctx();
    SA_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_fold_group), &g, sizeof(g)));
    tuning().fold_group = g;
}

namespace {

// Block sum of an NT-thread block (s_tmp: NT / kWave doubles), valid in
// thread 0.
template <int NT> __device__ __forceinline__ double block_sum(double v, double* s_tmp) {
    constexpr uint32_t nw = NT / kWave;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    const uint32_t w = threadIdx.x / kWave;
    if (lane_id() == 0) s_tmp[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (uint32_t k = 0; k < nw; ++k) t += s_tmp[k];
    __syncthreads();
    return t;
}

// Fold K per-thread values over the grid.  Every block stores its K block
// sums.  One level (c_fold_group = 0): the last block to take a ticket sums
// all partials in block order.  Two levels (c_fold_group = g): blocks take
// tickets on their group's counter (one 128-byte line per group), the last
// block of each group sums its group's partials in block order and takes a
// ticket on the global counter, and the last group sums the group totals in
// group order; no counter sees more than max(g, G/g) atomics, where the flat
// fold serialises G atomics on one address.  The last block returns true in
// all its threads with the totals in tot[]; counters are reset by the blocks
// that finish them.  Every order is fixed, so the totals are reproducible.
template <int K, int NT>
__device__ bool grid_fold(const double (&v)[K], double* partials, uint32_t* counter, double (&tot)[K]) {
    __shared__ double s_tmp[NT / kWave];
    __shared__ double s_tot[K];
    __shared__ bool s_last;
    const uint32_t G = gridDim.x;  // <= kMaxCgBlocks
    const uint32_t grp = c_fold_group;
    double bs[K];
#pragma unroll
    for (int k = 0; k < K; ++k) bs[k] = block_sum<NT>(v[k], s_tmp);
    constexpr int PER = kMaxCgBlocks / NT;
    if (grp == 0) {
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
#pragma unroll
        for (int k = 0; k < K; ++k) {
            // all loads issued before any add: one memory round trip
            double pv[PER];
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const uint32_t b = threadIdx.x + u * NT;
                pv[u] = b < G ? coherent_load(&partials[(size_t)k * G + b]) : 0.0;
            }
            double a = 0.0;
#pragma unroll
            for (int u = 0; u < PER; ++u) a += pv[u];
            const double t = block_sum<NT>(a, s_tmp);
            if (threadIdx.x == 0) s_tot[k] = t;
        }
        if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; ++k) tot[k] = s_tot[k];
        return true;
    }
    // two levels
    const uint32_t ng = (G + grp - 1) / grp;
    const uint32_t g = blockIdx.x / grp;
    const uint32_t gsize = min(grp, G - g * grp);
    uint32_t* gcount = counter + (size_t)(1 + g) * kCounterStride;
    double* gpart = partials + (size_t)K * kMaxCgBlocks;  // [K][ng] group totals
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) coherent_store(&partials[(size_t)k * G + blockIdx.x], bs[k]);
        vmem_release_order();
        const uint32_t t = __hip_atomic_fetch_add(gcount, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (t == gsize - 1);
    }
    __syncthreads();
    if (!s_last) return false;
    vmem_acquire_order();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t b = g * grp + threadIdx.x;
        const double a = threadIdx.x < gsize ? coherent_load(&partials[(size_t)k * G + b]) : 0.0;
        const double t = block_sum<NT>(a, s_tmp);
        if (threadIdx.x == 0) s_tot[k] = t;
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) coherent_store(&gpart[(size_t)k * ng + g], s_tot[k]);
        __hip_atomic_store(gcount, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        vmem_release_order();
        const uint32_t t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (t == ng - 1);
    }
    __syncthreads();
    if (!s_last) return false;
    vmem_acquire_order();
    constexpr int PER2 = (kMaxFoldGroups + NT - 1) / NT;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double pv[PER2];
#pragma unroll
        for (int u = 0; u < PER2; ++u) {
            const uint32_t q = threadIdx.x + u * NT;
            pv[u] = q < ng ? coherent_load(&gpart[(size_t)k * ng + q]) : 0.0;
        }
        double a = 0.0;
#pragma unroll
        for (int u = 0; u < PER2; ++u) a += pv[u];
        const double t = block_sum<NT>(a, s_tmp);
        if (threadIdx.x == 0) s_tot[k] = t;
    }
    if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) tot[k] = s_tot[k];
    return true;
}

// Block partials without a grid fold: each block stores its K block sums
// at p[k * stride + blockIdx.x]; the next kernel sums them (sum_partials).
// The producer ends with no ticket atomics and no last-block tail.  The
// wave sums of all K values cross LDS behind one barrier.
template <int K, int NT>
__device__ void store_partials(const double (&v)[K], double* __restrict__ p, uint32_t stride) {
    constexpr uint32_t nw = NT / kWave;
    __shared__ double s_w[K][nw];
    double a[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        a[k] = v[k];
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) a[k] += __shfl_xor(a[k], o, kWave);
    }
    const uint32_t wv = threadIdx.x / kWave;
    if (lane_id() == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) s_w[k][wv] = a[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double t = 0.0;
            for (uint32_t q = 0; q < nw; ++q) t += s_w[k][q];
            p[(size_t)k * stride + blockIdx.x] = t;
        }
}

// The solver's stop flag as a plain vector load (relaxed, agent scope), so
// that a kernel can issue its first data loads before waiting for it.
__device__ __forceinline__ uint32_t flag_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The K x G partials a previous kernel stored with store_partials, read as
// one index space e = k G + b: thread t adds entries t, t + NT, ...
// of every k in ascending order, then the lanes' sums meet in an xor
// butterfly (every lane gets the same bits, additions being commutative)
// and the waves' in LDS behind one barrier, in wave order.  Every block
// that sums the same partials therefore gets the same bits, so each block
// of a kernel can derive the same scalar (alpha, beta) without a grid-wide
// fold.
template <int K> __device__ __forceinline__ double partial_at(const double* p, uint32_t stride, uint32_t G,
                                                              uint32_t e) {
    uint32_t kk = 0;
#pragma unroll
    for (int k = 1; k < K; ++k) kk += e >= k * G ? 1u : 0u;
    return p[(size_t)kk * stride + (e - kk * G)];
}
template <int K> __device__ __forceinline__ void add_partial(double (&a)[K], uint32_t G, uint32_t e, double v) {
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (e >= k * G && e < (k + 1) * G) a[k] += v;
}
template <int K, int NT> __device__ void finish_partials(double (&a)[K], double (&tot)[K]) {
    constexpr uint32_t nw = NT / kWave;
    __shared__ double s_w[K][nw];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) a[k] += __shfl_xor(a[k], o, kWave);
    const uint32_t wv = threadIdx.x / kWave;
    if (lane_id() == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) s_w[k][wv] = a[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double t = 0.0;
        for (uint32_t q = 0; q < nw; ++q) t += s_w[k][q];
        tot[k] = t;
    }
}

// Sum of the K x G partials (G <= kMaxCgBlocks), LOADS loads per thread in
// flight per round.  Result in every thread.
template <int K, int NT, int LOADS = 16>
__device__ void sum_partials(const double* __restrict__ p, uint32_t stride, uint32_t G, double (&tot)[K]) {
    double a[K];
#pragma unroll
    for (int k = 0; k < K; ++k) a[k] = 0.0;
    const uint32_t total = K * G;
    for (uint32_t e0 = 0; e0 < total; e0 += LOADS * NT) {
        double pv[LOADS];
#pragma unroll
        for (int u = 0; u < LOADS; ++u) {
            const uint32_t e = e0 + threadIdx.x + u * NT;
            pv[u] = e < total ? partial_at<K>(p, stride, G, e) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < LOADS; ++u) add_partial<K>(a, G, e0 + threadIdx.x + u * NT, pv[u]);
    }
    finish_partials<K, NT>(a, tot);
}

// The same sum in two halves: the loads issued early (prefetch_partials,
// when K x G <= kPrefetch NT, kPrefetch doubles per thread), the adds after
// other work (sum_prefetched); larger grids take sum_partials at the
// point of use, two loads per round so that the caller's registers stay
// few.  Both orders give sum_partials' bits.
constexpr int kPrefetch = 4;
struct PartialsPrefetch {
    double v[kPrefetch];
    bool whole;
};
template <int K, int NT>
__device__ __forceinline__ PartialsPrefetch prefetch_partials(const double* __restrict__ p, uint32_t stride,
                                                              uint32_t G) {
    PartialsPrefetch f;
    const uint32_t total = K * G;
    f.whole = total <= kPrefetch * NT;
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
        const uint32_t e = threadIdx.x + u * NT;
        f.v[u] = (f.whole && e < total) ? partial_at<K>(p, stride, G, e) : 0.0;
    }
    return f;
}
template <int K, int NT>
__device__ void sum_prefetched(const PartialsPrefetch& f, const double* __restrict__ p, uint32_t stride, uint32_t G,
                               double (&tot)[K]) {
    if (!f.whole) {
        sum_partials<K, NT, 2>(p, stride, G, tot);
        return;
    }
    double a[K];
#pragma unroll
    for (int k = 0; k < K; ++k) a[k] = 0.0;
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
        const uint32_t e = threadIdx.x + u * NT;
        if (e < K * G) add_partial<K>(a, G, e, f.v[u]);
    }
    finish_partials<K, NT>(a, tot);
}



// rows of an NT-thread kernel, grid stride
#define CG_ROWS(i) for (uint32_t i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT)

// rows of an NT-thread kernel taken by groups of GS lanes (GS divides the
// wavefront, so a group never straddles one); g = lane within the group
#define CG_GROUP_ROWS(i, g, GS)                                                                   \
    for (uint32_t _t = blockIdx.x * NT + threadIdx.x, g = _t % (GS), i = _t / (GS); i < n;       \
         i += gridDim.x * (NT / (GS)))

// FP64 (A x)_i of a CSR row by the GS lanes of a group: lane g adds entries
// g, g + GS, ... in order, an xor butterfly combines the lanes (every lane
// gets the same bits).  The lanes read consecutive entries, so a row's
// columns and values arrive in one or two coalesced loads instead of one
// load per entry from a lone lane.
template <uint32_t GS>
__device__ __forceinline__ double group_row_dot(const uint32_t* __restrict__ row, const uint32_t* __restrict__ col,
                                                const double* __restrict__ val, const double* __restrict__ x,
                                                uint32_t i, uint32_t g) {
    double s = 0.0;
    const uint32_t e = row[i + 1];
    for (uint32_t k = row[i] + g; k < e; k += GS) s += val[k] * x[col[k]];
#pragma unroll
    for (uint32_t o = GS / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, GS);
    return s;
}

template <class V> struct V2T;
template <> struct V2T<float> {
    using T = float2;
};
template <> struct V2T<double> {
    using T = double2;
};
template <class V> using V2 = typename V2T<V>::T;

// ---------------------------------------------------------------------------
// SELL-64 construction
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool in_part(uint32_t c, uint32_t i, int part) {
    return part == kSellOffDiag ? c != i : (part == kSellLower ? c < i : c > i);
}

// One wave per slice: row lengths, the slice width (max length) and the
// slice's entry count (its number of slots).
__global__ __launch_bounds__(256) void sell_len_kernel(uint32_t n, uint32_t nslices, const uint32_t* __restrict__ row,
                                                       const uint32_t* __restrict__ col, int part,
                                                       uint16_t* __restrict__ len, uint16_t* __restrict__ width,
                                                       uint32_t* __restrict__ count) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t s = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (s >= nslices) return;
    const uint32_t i = s * kSlice + lane;
    uint32_t l = 0;
    if (i < n)
        for (uint32_t e = row[i]; e < row[i + 1]; ++e) l += in_part(col[e], i, part) ? 1u : 0u;
    if (i < n) len[i] = (uint16_t)l;
    uint32_t m = l, tot = l;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
        tot += (uint32_t)__shfl_xor((int)tot, o, 64);
    }
    if (lane == 0) {
        width[s] = (uint16_t)m;
        count[s] = tot;
    }
}

// Slot of this lane's entry at a step whose participating lanes are the set
// bits of m (the step starts at slot seg).
__device__ __forceinline__ uint32_t step_slot(uint32_t seg, uint64_t m) {
    return seg + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

constexpr uint32_t kFillSteps = 256;  // step tables held in LDS per slice

// One block of 256 threads per slice, 4 lanes per row: a wavefront reads
// 16 rows' entries 4 at a time.  Wavefront 0 first records, for each step k
// of the slice, the ballot of the rows longer than k and the step's first
// slot (in LDS, up to kFillSteps steps).  The k-th part entry of a row
// (k from the ballot of the row's 4 lanes) goes to step k, at the step's
// first slot + the number of lower rows in the ballot; steps beyond
// kFillSteps (rows that long are rare) count the slice's lengths directly.
// The diagonal is the row's diagonal entry (standard form) or its FP64 row
// sum (row-sum form), added entry by entry in row order through the row's
// lanes, as a one-lane walk of the row would.
constexpr uint32_t kFillGroup = 4;
static_assert(kSlice * kFillGroup == 256, "a block of 256 threads holds one slice");
template <class V>
__global__ __launch_bounds__(256) void sell_fill_kernel(uint32_t n, uint32_t nslices,
                                                        const uint32_t* __restrict__ row,
                                                        const uint32_t* __restrict__ col,
                                                        const double* __restrict__ val64,
                                                        const float* __restrict__ val32, int form, int part,
                                                        const uint32_t* __restrict__ off,
                                                        const uint16_t* __restrict__ width,
                                                        const uint16_t* __restrict__ len,
                                                        uint32_t* __restrict__ scol, int16_t* __restrict__ scol16,
                                                        V* __restrict__ sval, V* __restrict__ diag,
                                                        uint32_t* __restrict__ entry_slot) {
    __shared__ uint64_t smask[kFillSteps];
    __shared__ uint32_t sfirst[kFillSteps];
    const uint32_t s = blockIdx.x;
    if (s >= nslices) return;  // the whole block
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t base = off[s];
    const uint32_t W = width[s];
    if (threadIdx.x < 64) {
        const uint32_t i = s * kSlice + lane;
        const uint32_t L = i < n ? (uint32_t)len[i] : 0u;
        uint32_t seg = base;
        const uint32_t ws = W < kFillSteps ? W : kFillSteps;
        for (uint32_t k = 0; k < ws; ++k) {
            const uint64_t m = __ballot(k < L);
            if (lane == 0) {
                smask[k] = m;
                sfirst[k] = seg;
            }
            seg += (uint32_t)__popcll(m);
        }
    }
    __syncthreads();
    const uint32_t rl = threadIdx.x / kFillGroup;  // row within the slice
    const uint32_t sub = threadIdx.x % kFillGroup;
    const uint32_t i = s * kSlice + rl;
    const bool live = i < n;
    const uint32_t e0 = live ? row[i] : 0u, e1 = live ? row[i + 1] : 0u;
    const uint32_t gshift = lane & ~(kFillGroup - 1u);  // the row's lanes in a ballot
    const uint64_t below_row = (1ull << rl) - 1ull;
    uint32_t kbase = 0;
    bool has_diag = false;
    double rowsum = 0.0;
    for (uint32_t q = 0;; q += kFillGroup) {
        const uint32_t e = e0 + q + sub;
        const bool act = e < e1;
        if (__ballot(act) == 0) break;
        const uint32_t c = act ? col[e] : 0u;
        const double v = act ? (val64 ? val64[e] : (double)val32[e]) : 0.0;
        if (form == kSellRowSum) {
#pragma unroll
            for (uint32_t j = 0; j < kFillGroup; ++j) {
                const double vj = __shfl(v, (int)(gshift + j), 64);
                if (e0 + q + j < e1) rowsum += vj;
            }
        }
        const bool inp = act && in_part(c, i, part);
        const uint32_t grp = (uint32_t)(__ballot(inp) >> gshift) & ((1u << kFillGroup) - 1u);
        if (inp) {
            const uint32_t k = kbase + (uint32_t)__popc(grp & ((1u << sub) - 1u));
            uint32_t slot;
            if (k < kFillSteps) {
                slot = sfirst[k] + (uint32_t)__popcll(smask[k] & below_row);
            } else {
                slot = base;
                for (uint32_t l = 0; l < kSlice; ++l) {
                    const uint32_t r = s * kSlice + l;
                    const uint32_t Ll = r < n ? (uint32_t)len[r] : 0u;
                    slot += (Ll < k ? Ll : k) + ((l < rl && Ll > k) ? 1u : 0u);
                }
            }
            const int32_t o = (int32_t)c - (int32_t)i;
            scol[slot] = c;
            scol16[slot] = (o >= -32767 && o <= 32767) ? (int16_t)o : kColEscape;
            sval[slot] = (V)v;
            if (entry_slot) entry_slot[e] = slot;
        }
        const bool isd = act && c == i;
        if (isd && form != kSellRowSum) diag[i] = (V)v;
        has_diag |= ((uint32_t)(__ballot(isd) >> gshift) & ((1u << kFillGroup) - 1u)) != 0u;
        kbase += (uint32_t)__popc(grp);
    }
    if (live && sub == 0) {
        if (form == kSellRowSum)
            diag[i] = (V)rowsum;
        else if (!has_diag)
            diag[i] = (V)0;  // a row without a diagonal entry
    }
}

}  // namespace

template <class V>
void build_sell(uint32_t n, const uint32_t* row, const uint32_t* col, const double* val64, const float* val32,
                SellForm form, SellPart part, Sell<V>& S, hipStream_t s, uint32_t* entry_slot) {
    S.n = n;
    S.nslices = (n + kSlice - 1) / kSlice;
    const uint32_t ns = S.nslices ? S.nslices : 1;
    S.off.resize(ns + 1);
    S.width.resize(ns);
    S.len.resize(n ? n : 1);
    S.diag.resize(n ? n : 1);
    S.tmp.resize(ns);
    S.scan_ws.resize(scan_workspace_words(ns));
    if (!n) {
        S.entries = 0;
        return;
    }
    hipLaunchKernelGGL(sell_len_kernel, dim3((S.nslices * 64 + 255) / 256), dim3(256), 0, s, n, S.nslices, row, col,
                       (int)part, S.len.p, S.width.p, S.tmp.p);
    exclusive_scan_u32(S.tmp.p, S.off.p, S.nslices, S.scan_ws.p, s);
    S.entries = read_scalar(S.off.p + S.nslices);
    check_error_flag();  // a scan look-back that ran out of spins
    // kSlice spare slots: the slot a lane computes for a step it sits out
    // can be one past the last entry
    S.col.resize(S.entries + kSlice);
    S.col16.resize(S.entries + kSlice);
    S.val.resize(S.entries + kSlice);
    hipLaunchKernelGGL((sell_fill_kernel<V>), dim3(S.nslices), dim3(256), 0, s, n, S.nslices, row, col, val64, val32,
                       (int)form, (int)part, S.off.p, S.width.p, S.len.p, S.col.p, S.col16.p, S.val.p, S.diag.p,
                       entry_slot);
}
template void build_sell<float>(uint32_t, const uint32_t*, const uint32_t*, const double*, const float*, SellForm,
                                SellPart, Sell<float>&, hipStream_t, uint32_t*);
template void build_sell<double>(uint32_t, const uint32_t*, const uint32_t*, const double*, const float*, SellForm,
                                 SellPart, Sell<double>&, hipStream_t, uint32_t*);

namespace {

// Matrix arguments of the SELL kernels.  map selects how slices are dealt
// to wavefronts (slice_span).
template <class V> struct SellArgs {
    uint32_t n, nslices;
    const uint32_t* off;
    const uint16_t* width;
    const uint16_t* len;
    const uint32_t* col;
    const int16_t* col16;
    const V* val;
    const V* diag;
    uint32_t map;
};
template <class V> SellArgs<V> sell_args(const Sell<V>& S) {
    return {S.n, S.nslices, S.off.p, S.width.p, S.len.p, S.col.p, S.col16.p, S.val.p, S.diag.p, tuning().xcd_map};
}

// Entries k .. k + CH - 1 of a row (row length L, row ii) from the steps
// starting at slot seg (advanced past them): values and columns, past the
// row's end column ii with value 0.  Every lane of the wavefront takes part
// (the ballots).  The 16-bit offsets and the values are all loaded before
// any escaped column (one entry in a few hundred on Morton-ordered meshes)
// is fetched from the 32-bit copy.
template <class V, int CH>
__device__ __forceinline__ void sell_load(const SellArgs<V>& M, uint32_t& seg, uint32_t k, uint32_t L, uint32_t ii,
                                          uint32_t (&c)[CH], V (&a)[CH]) {
    // offsets widened to 32 bits as they load: two 16-bit values packed in
    // one register would need each load to land before the next one issues
    int32_t o[CH];
    uint32_t e[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
        const bool ok = k + u < L;
        const uint64_t m = __ballot(ok);
        e[u] = step_slot(seg, m);
        seg += (uint32_t)__popcll(m);
        o[u] = ok ? (int32_t)M.col16[e[u]] : 0;
        a[u] = ok ? M.val[e[u]] : (V)0;
    }
#pragma unroll
    for (int u = 0; u < CH; ++u)
        c[u] = o[u] == (int32_t)kColEscape ? M.col[e[u]] : (uint32_t)((int32_t)ii + o[u]);
}

// Slices of one wavefront: [s, e) with stride step.
//  map 0: grid stride (wave w takes w, w + W, ...);
//  map 1: contiguous chunks per wave with the blocks renumbered so that the
//         blocks one XCD receives (hardware deals workgroups round-robin to
//         the 8 XCDs: block b runs on XCD b % 8) own one contiguous eighth
//         of the rows.  Rows are in Morton order, so the neighbour entries a
//         wave gathers were mostly fetched into the same XCD's L2 by the
//         neighbouring waves.
constexpr uint32_t kXcds = 8;
struct SliceSpan {
    uint32_t s, e, step;
};
__device__ __forceinline__ SliceSpan slice_span(uint32_t nslices, uint32_t map) {
    const uint32_t wpb = blockDim.x >> 6;
    const uint32_t G = gridDim.x;
    const uint32_t nw = G * wpb;
    if (map == 0) return {blockIdx.x * wpb + (threadIdx.x >> 6), nslices, nw};
    const uint32_t b = blockIdx.x;
    const uint32_t lb = (G % kXcds == 0) ? (b % kXcds) * (G / kXcds) + b / kXcds : b;
    const uint32_t gw = lb * wpb + (threadIdx.x >> 6);
    const uint32_t chunk = (nslices + nw - 1) / nw;
    const uint32_t s0 = min(gw * chunk, nslices);
    return {s0, min(s0 + chunk, nslices), 1u};
}

// R slices of one wavefront processed together (s, s + step, ...): this
// lane's row in each, its length, the slice's first slot, and the widest
// slice.
// R slices at once put R rows' loads in flight before any of their
// gathers; the solvers instantiate R = 1, since R = 2 measured slower on
// the 1M systems (more registers, fewer resident waves, no shorter chain).
template <int R> struct SliceRows {
    uint32_t i[R], ii[R], L[R], base[R];
    bool valid[R];
    uint32_t W;
};
template <class V, int R>
__device__ __forceinline__ SliceRows<R> slice_rows(const SellArgs<V>& M, uint32_t s, const SliceSpan& sp,
                                                   uint32_t lane) {
    SliceRows<R> q;
    q.W = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t sr = s + r * sp.step;
        const bool has = sr < sp.e;
        q.i[r] = sr * kSlice + lane;
        q.valid[r] = has && q.i[r] < M.n;
        q.ii[r] = q.valid[r] ? q.i[r] : 0u;
        q.L[r] = q.valid[r] ? (uint32_t)M.len[q.i[r]] : 0u;
        q.base[r] = has ? M.off[sr] : 0u;
        q.W = max(q.W, has ? (uint32_t)M.width[sr] : 0u);
    }
    return q;
}

// acc[r] = term(r, acc[r], a_k, fetch(c_k)) over the entries k of this
// lane's R rows in ascending order (steps past a row's length read as
// column ii, value 0).  Per step the CH entries of all R rows are loaded,
// then all their gathers are issued, then the terms are added.
template <class V, int R, int CH, class G, class Fetch, class Term>
__device__ __forceinline__ void sell_rows_product(const SellArgs<V>& M, const SliceRows<R>& q, V (&acc)[R],
                                                  Fetch fetch, Term term) {
    uint32_t seg[R];
#pragma unroll
    for (int r = 0; r < R; ++r) seg[r] = q.base[r];
    for (uint32_t k = 0; k < q.W; k += CH) {
        uint32_t c[R][CH];
        V a[R][CH];
#pragma unroll
        for (int r = 0; r < R; ++r) sell_load<V, CH>(M, seg[r], k, q.L[r], q.ii[r], c[r], a[r]);
        G g[R][CH];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int u = 0; u < CH; ++u) g[r][u] = fetch(c[r][u]);
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int u = 0; u < CH; ++u) acc[r] = term(r, acc[r], a[r][u], g[r][u]);
    }
}

// Loop over the slices of this wavefront, R at a time (the kernel's matrix
// is M; q is the SliceRows of the current group).
#define SELL_FOR_SLICE_ROWS(R, q)                                                                    \
    for (SliceSpan _span = slice_span(M.nslices, M.map), _once = {0, 1, 1}; _once.s < _once.e;    \
         ++_once.s)                                                                                  \
        for (uint32_t _s = _span.s; _s < _span.e; _s += (R) * _span.step)                          \
            if (const SliceRows<R> q = slice_rows<V, R>(M, _s, _span, lane); true)

// ---------------------------------------------------------------------------
// init / replacement: r = b - A64 x64 (FP64), stored in precision V;
// Jacobi / no preconditioner: z into the current {z, d} buffer (P is not
// touched on the L L^T path); partial r.r, r.z, b.b.
// mode 0: init (d = 0, beta = 0, iter = 0)
// mode 1: replacement of the recursive residual by the true one.  The
//         direction d and beta (computed by the recursion) are kept and only
//         r, z and r.z change: when the recursive residual had drifted far
//         below the true one, this degrades to a restart instead of scaling
//         the old direction by r_true.z / r_rec.z.
// ---------------------------------------------------------------------------
template <class V, int PREC, int NT, uint32_t GS>
__global__ __launch_bounds__(NT) void cg_residual_kernel(
    uint32_t n, const uint32_t* __restrict__ row, const uint32_t* __restrict__ col, const double* __restrict__ val,
    const double* __restrict__ b, const double* __restrict__ x, V* __restrict__ r, uint32_t rs,
    V2<V>* __restrict__ P, const V* __restrict__ idiag, CgScalars* sc, double* partials, uint32_t* counter, int mode,
    double rtol2) {
    double v[3] = {0.0, 0.0, 0.0};
    CG_GROUP_ROWS(i, g, GS) {
        const double ax = group_row_dot<GS>(row, col, val, x, i, g);
        if (g != 0) continue;
        const double ri = b[i] - ax;
        const V rv = (V)ri;
        r[(size_t)i * rs] = rv;
        if constexpr (PREC != kPrecLLt) {
            V2<V> p = P[i];
            if (mode == 0) p.y = (V)0;
            const V zv = (PREC == kPrecJacobi) ? (V)(idiag[i] * rv) : rv;
            p.x = zv;
            P[i] = p;
            v[1] += (double)rv * (double)zv;
        }
        v[0] += (double)rv * (double)rv;
        v[2] += b[i] * b[i];
    }
    double tot[3];
    if (!grid_fold<3, NT>(v, partials, counter, tot)) return;
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

// d = z + beta d_old; q = A d; the block partials of d.q go to the update
// kernel, which derives alpha from them.
// DIFF: the values are in row-sum form (off-diagonals K_ij, diagonal slot
// holding the row sum s_i) and q_i = s_i d_i + sum_j K_ij (d_j - d_i).  For
// a graph Laplacian the differences of neighbouring entries of a smooth
// vector are exact in floating point (Sterbenz), so the FP32 product keeps
// its relative accuracy on the smooth modes that standard-form summation
// loses to cancellation.
template <class V, bool DIFF, int CH, int R, int NT>
__global__ __launch_bounds__(NT) void cg_dir_kernel(SellArgs<V> M, const V2<V>* __restrict__ cur,
                                                            V2<V>* __restrict__ nxt, V* __restrict__ q, uint32_t qs,
                                                            const CgScalars* sc, double* partials) {
    if (sc->done) return;
    const V beta = (V)sc->beta;
    const uint32_t lane = threadIdx.x & 63u;
    double v[1] = {0.0};
    SELL_FOR_SLICE_ROWS(R, rows) {
        V dn[R], acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const V2<V> p = cur[rows.ii[r]];
            dn[r] = p.x + beta * p.y;
            acc[r] = M.diag[rows.ii[r]] * dn[r];
        }
        sell_rows_product<V, R, CH, V2<V>>(
            M, rows, acc, [&](uint32_t c) { return cur[c]; },
            [&](int r, V ac, V a, V2<V> g) {
                const V gd = g.x + beta * g.y;
                if constexpr (DIFF)
                    return ac + a * (gd - dn[r]);
                else
                    return ac + a * gd;
            });
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (rows.valid[r]) {
                const uint32_t i = rows.i[r];
                nxt[i].y = dn[r];
                q[(size_t)i * qs] = acc[r];
                v[0] += (double)dn[r] * (double)acc[r];
            }
    }
    store_partials<1, NT>(v, partials + kJacobiDqSlot, kMaxCgBlocks);
}

// Direction kernel of the L L^T path: d is already formed (by the lower
// apply of the previous iteration, which knows beta), so the row loop
// gathers 4-byte d_j instead of {z, d} pairs and writes nothing but q.
// q = A d (row-sum form when DIFF), q into the {r, q} pair; the block
// partials of d.q go to partials[0][block] and the update kernel derives
// alpha from them.
template <class V, bool DIFF, int CH, int R, int NT>
__global__ __launch_bounds__(NT) void llt_dir_kernel(SellArgs<V> M, const V* __restrict__ d,
                                                             V* __restrict__ q, uint32_t qs, const CgScalars* sc,
                                                             double* partials) {
    // the stop flag is read with the first slice's loads in flight and
    // tested before anything is stored (every block sees the same flag)
    const uint32_t done = flag_load(&sc->done);
    const uint32_t lane = threadIdx.x & 63u;
    double v[1] = {0.0};
    SELL_FOR_SLICE_ROWS(R, rows) {
        V dn[R], acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            dn[r] = d[rows.ii[r]];
            acc[r] = M.diag[rows.ii[r]] * dn[r];
        }
        if (done) return;
        sell_rows_product<V, R, CH, V>(
            M, rows, acc, [&](uint32_t c) { return d[c]; },
            [&](int r, V ac, V a, V g) {
                if constexpr (DIFF)
                    return ac + a * (g - dn[r]);
                else
                    return ac + a * g;
            });
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (rows.valid[r]) {
                q[(size_t)rows.i[r] * qs] = acc[r];
                v[0] += (double)dn[r] * (double)acc[r];
            }
    }
    if (done) return;
    store_partials<1, NT>(v, partials, kMaxCgBlocks);
}

// x += alpha d; r -= alpha q; z = M r (Jacobi / none) into nxt.x;
// partial r.z, r.r
template <class V, class XV, int PREC, int NT>
__global__ __launch_bounds__(NT) void cg_update_kernel(uint32_t n, XV* __restrict__ x, V* __restrict__ r,
                                                               V2<V>* __restrict__ nxt, const V* __restrict__ q,
                                                               const V* __restrict__ idiag, CgScalars* sc,
                                                               double* partials, uint32_t* counter) {
    if (sc->done) return;
    // alpha = r.z / d.q from the direction kernel's block partials, summed
    // by every block (block 0 records it)
    double dq[1];
    sum_partials<1, NT>(partials + kJacobiDqSlot, kMaxCgBlocks, gridDim.x, dq);
    const double alpha = sc->rz / dq[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->dq = dq[0];
        sc->alpha = alpha;
    }
    const V av = (V)alpha;
    double v[2] = {0.0, 0.0};
    CG_ROWS(i) {
        V2<V> p = nxt[i];
        x[i] += (XV)(alpha * (double)p.y);
        const V rv = r[i] - av * q[i];
        r[i] = rv;
        v[1] += (double)rv * (double)rv;
        if constexpr (PREC != kPrecLLt) {
            const V zv = (PREC == kPrecJacobi) ? idiag[i] * rv : rv;
            p.x = zv;
            nxt[i] = p;
            v[0] += (double)rv * (double)zv;
        }
    }
    double tot[2];
    if (!grid_fold<2, NT>(v, partials, counter, tot)) return;
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

// L L^T path, one kernel for the update and the first half of the apply:
// x += alpha d; r' = r - alpha q; w = L^T r' (w_i = L_ii r'_i + sum_{j>i}
// L_ji r'_j, U = the upper SELL).  r and q live interleaved ({r, q} per
// row, written by the direction kernel), so the r' of each upper neighbour
// is recomputed from one 8-byte load with the same operations its owner
// uses; r' goes to the other {r, q} buffer.  Every block derives alpha =
// r.z / d.q from the direction kernel's partials (block 0 records alpha,
// d.q and r.z as rz_old); the block partials of w.w (= r'.z, since
// z = L w) and r'.r' go to partials[1..2][block] for the lower apply.
// mode 0: iteration; mode 1: init, mode 2: after a residual replacement
// (r given in rq.x, alpha = 0, x and rq_nxt untouched).
template <class V, class XV, int CH, int R, int NT>
__global__ __launch_bounds__(NT) void llt_update_kernel(SellArgs<V> U, XV* __restrict__ x,
                                                                const V* __restrict__ d,
                                                                const V2<V>* __restrict__ rq,
                                                                V2<V>* __restrict__ rq_nxt, V* __restrict__ w,
                                                                CgScalars* sc, double* partials, int mode) {
    // the stop flag is read with the first loads in flight and tested
    // before the partials' sum (a barrier) and before anything is stored
    const uint32_t done = mode == 0 ? flag_load(&sc->done) : 0u;
    const SellArgs<V>& M = U;
    const bool it = mode == 0;
    // the direction kernel's partials and the first slice group's own data
    // are loaded before the partials' sum (alpha) is taken
    PartialsPrefetch pf{};
    if (it) pf = prefetch_partials<1, NT>(partials, kMaxCgBlocks, gridDim.x);
    const uint32_t lane = threadIdx.x & 63u;
    const SliceSpan sp = slice_span(M.nslices, M.map);
    const SliceRows<R> first = slice_rows<V, R>(M, sp.s, sp, lane);
    V2<V> own0[R];
    V diag0[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        own0[r] = rq[first.ii[r]];
        diag0[r] = U.diag[first.ii[r]];
    }
    double alpha = 0.0;
    if (done) return;
    if (it) {
        double dq[1];
        sum_prefetched<1, NT>(pf, partials, kMaxCgBlocks, gridDim.x, dq);
        const double rz = sc->rz;
        alpha = rz / dq[0];
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            sc->dq = dq[0];
            sc->alpha = alpha;
            sc->rz_old = rz;
        }
    }
    const V av = (V)alpha;
    double v[2] = {0.0, 0.0};
    auto body = [&](const SliceRows<R>& rows, const V2<V> (&own)[R], const V (&dg)[R]) {
        V ri[R], acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            ri[r] = it ? own[r].x - av * own[r].y : own[r].x;
            acc[r] = dg[r] * ri[r];
        }
        sell_rows_product<V, R, CH, V2<V>>(
            U, rows, acc, [&](uint32_t c) { return rq[c]; },
            [&](int, V ac, V a, V2<V> g) { return ac + a * (it ? g.x - av * g.y : g.x); });
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (rows.valid[r]) {
                const uint32_t i = rows.i[r];
                if (it) {
                    x[i] += (XV)(alpha * (double)d[i]);
                    rq_nxt[i].x = ri[r];
                }
                w[i] = acc[r];
                v[0] += (double)acc[r] * (double)acc[r];
                v[1] += (double)ri[r] * (double)ri[r];
            }
    };
    body(first, own0, diag0);
    for (uint32_t s0 = sp.s + R * sp.step; s0 < sp.e; s0 += R * sp.step) {
        const SliceRows<R> rows = slice_rows<V, R>(M, s0, sp, lane);
        V2<V> own[R];
        V dg[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            own[r] = rq[rows.ii[r]];
            dg[r] = U.diag[rows.ii[r]];
        }
        body(rows, own, dg);
    }
    store_partials<2, NT>(v, partials + kMaxCgBlocks, kMaxCgBlocks);
}

// Second half of the apply, z = L w (z_i = L_ii w_i + sum_{j<i} L_ij w_j),
// and the next direction d_new = z + beta d_old; z is never stored.  Every
// block sums the update kernel's partials: mode 0 (iteration) beta =
// (r'.z) / (r.z) with r.z = rz_old, and block 0 records r.z, r.r, beta,
// the iteration count and the stop flag; mode 1 (init) d_new = z (d_old
// not read), block 0 records r.z and the stop flag; mode 2 (after a
// residual replacement) beta is the recursion's, as in cg_residual_kernel,
// and block 0 records r.z.  mode 0 is skipped once the solve is flagged
// done; the iteration that sets the flag may leave d_new partly written
// (blocks that start after block 0 wrote it return), which is harmless:
// a residual replacement rebuilds that buffer from d_old before it is read.
template <class V, int CH, int R, int NT>
__global__ __launch_bounds__(NT) void llt_lo_kernel(SellArgs<V> M, const V* __restrict__ w,
                                                               const V* __restrict__ d_old, V* __restrict__ d_new,
                                                               CgScalars* sc, const double* partials, int mode) {
    // the stop flag is read with the first loads in flight and tested
    // before the partials' sum (a barrier) and before anything is stored
    const uint32_t done = mode == 0 ? flag_load(&sc->done) : 0u;
    // the partials' loads go out first; their sum (beta) is taken after the
    // wavefront's first slice group, whose loads cover their latency
    const PartialsPrefetch pf = prefetch_partials<2, NT>(partials + kMaxCgBlocks, kMaxCgBlocks, gridDim.x);
    const uint32_t lane = threadIdx.x & 63u;
    const SliceSpan sp = slice_span(M.nslices, M.map);
    auto product = [&](const SliceRows<R>& rows, V (&acc)[R]) {
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = M.diag[rows.ii[r]] * w[rows.ii[r]];
        sell_rows_product<V, R, CH, V>(
            M, rows, acc, [&](uint32_t c) { return w[c]; }, [&](int, V ac, V a, V g) { return ac + a * g; });
    };
    auto store = [&](const SliceRows<R>& rows, const V (&acc)[R], V beta) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (rows.valid[r]) {
                const uint32_t i = rows.i[r];
                d_new[i] = mode == 1 ? acc[r] : acc[r] + beta * d_old[i];
            }
    };
    // first group (empty when the wavefront has no slices: no rows valid)
    const SliceRows<R> first = slice_rows<V, R>(M, sp.s, sp, lane);
    V acc0[R];
    product(first, acc0);
    if (done) return;
    double tot[2];
    sum_prefetched<2, NT>(pf, partials + kMaxCgBlocks, kMaxCgBlocks, gridDim.x, tot);
    double bd = 0.0;
    if (mode == 0)
        bd = tot[0] / sc->rz_old;
    else if (mode == 2)
        bd = sc->beta;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->rz = tot[0];
        if (mode == 0) {
            sc->rr = tot[1];
            sc->beta = bd;
            sc->iter += 1;
            const double cv = sc->crit ? tot[0] : tot[1];
            sc->done = (cv <= sc->tol2 || sc->iter >= sc->maxit) ? 1u : 0u;
        } else if (mode == 1) {
            sc->rz_old = tot[0];
            const double cv = sc->crit ? tot[0] : sc->rr;
            sc->done = (cv <= sc->tol2 || sc->iter >= sc->maxit) ? 1u : 0u;
        }
    }
    const V beta = (V)bd;
    store(first, acc0, beta);
    for (uint32_t s0 = sp.s + R * sp.step; s0 < sp.e; s0 += R * sp.step) {
        const SliceRows<R> rows = slice_rows<V, R>(M, s0, sp, lane);
        V acc[R];
        product(rows, acc);
        store(rows, acc, beta);
    }
}

template <int NT, uint32_t GS>
__global__ __launch_bounds__(NT) void true_residual_kernel(uint32_t n, const uint32_t* __restrict__ row,
                                                                   const uint32_t* __restrict__ col,
                                                                   const double* __restrict__ val,
                                                                   const double* __restrict__ b,
                                                                   const double* __restrict__ x, double* partials,
                                                                   uint32_t* counter, double* out2) {
    double v[2] = {0.0, 0.0};
    CG_GROUP_ROWS(i, g, GS) {
        const double ax = group_row_dot<GS>(row, col, val, x, i, g);
        if (g != 0) continue;
        const double ri = b[i] - ax;
        v[0] += ri * ri;
        v[1] += b[i] * b[i];
    }
    double tot[2];
    if (!grid_fold<2, NT>(v, partials, counter, tot)) return;
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

// Vectors of one solve, carved from the grow-only workspace.
template <class V> struct CgVecs {
    V2<V>* P[2];   // {z, d} ping-pong (Jacobi / no preconditioner)
    V2<V>* RQ[2];  // {r, q} ping-pong (L L^T path)
    V* D[2];       // d ping-pong (L L^T path)
    V* r;
    V* q;
    V* w;
};
template <class V> CgVecs<V> carve(CgWork& w, uint32_t n) {
    const size_t nv = ((size_t)n + 63) & ~(size_t)63;
    w.vbuf.resize(nv * (4 * sizeof(V2<V>) + 5 * sizeof(V)) + 256);
    uint8_t* p = w.vbuf.p;
    CgVecs<V> v;
    v.P[0] = reinterpret_cast<V2<V>*>(p);
    p += nv * sizeof(V2<V>);
    v.P[1] = reinterpret_cast<V2<V>*>(p);
    p += nv * sizeof(V2<V>);
    v.RQ[0] = reinterpret_cast<V2<V>*>(p);
    p += nv * sizeof(V2<V>);
    v.RQ[1] = reinterpret_cast<V2<V>*>(p);
    p += nv * sizeof(V2<V>);
    v.D[0] = reinterpret_cast<V*>(p);
    p += nv * sizeof(V);
    v.D[1] = reinterpret_cast<V*>(p);
    p += nv * sizeof(V);
    v.r = reinterpret_cast<V*>(p);
    p += nv * sizeof(V);
    v.q = reinterpret_cast<V*>(p);
    p += nv * sizeof(V);
    v.w = reinterpret_cast<V*>(p);
    p += nv * sizeof(V);

    return v;
}

// ---------------------------------------------------------------------------
// Driver.  V: iteration precision (matrix values, vectors).  The solution
// updates x += alpha d go to the FP64 x itself in every precision: an FP32
// correction merged at the residual replacements carries rounding noise that
// is rough on the mesh scale, which A amplifies against the smooth
// right-hand side (16% more iterations on the 1M dam-break).
// ---------------------------------------------------------------------------
template <class V, bool DIFF, int PREC, int CH, int R, int NT>
void cg_drive(const DevCsr<double>& A64, const Sell<V>& S, const V* idiag, const LowerFactor<V>* L, const double* b,
              double* x, const CgOptions& o, CgWork& w, SolveStats& st, hipStream_t s) {
    using XV = double;
    const uint32_t n = A64.n;
    w.resize(n, NT);
    const dim3 grid(w.blocks), blk(NT);
    CgVecs<V> vv = carve<V>(w, n);
    XV* xacc = x;
    const SellArgs<V> MA = sell_args(S);
    SellArgs<V> ML{}, MU{};
    if constexpr (PREC == kPrecLLt) {
        fill<V>(vv.D[1], n, (V)0, s);  // d_{-1}: defined values even where beta = 0
        ML = sell_args(L->lo);
        MU = sell_args(L->up);
        ML.diag = L->diag.p;
        MU.diag = L->diag.p;
    }
    hipEvent_t e0, e1;
    SA_HIP(hipEventCreate(&e0));
    SA_HIP(hipEventCreate(&e1));
    SA_HIP(hipEventRecord(e0, s));

    publish_options(w, o, s);
    const double rtol2 = o.abs_tol2 >= 0.0 ? -o.abs_tol2 : o.rtol * o.rtol;
    // L L^T buffers: iteration `it` reads d from D[it & 1] and r from
    // RQ[it & 1], and leaves d and r for it + 1 in the other halves; a
    // replacement after iteration it - 1 rebuilds D[it & 1] from the
    // direction D[(it + 1) & 1] it was made from.
    // FP64 residual rows by groups of 8 lanes (2D, ~7 entries per row) or
    // 16 (3D, ~15)
    const bool wide_rows = (uint64_t)A64.nnz > 8ull * n;
    auto residual_kernel = [&](V* rp, uint32_t rs, V2<V>* cur, int mode) {
        if (wide_rows)
            hipLaunchKernelGGL((cg_residual_kernel<V, PREC, NT, 16>), grid, blk, 0, s, n, A64.row.p, A64.col.p,
                               A64.val.p, b, x, rp, rs, cur, idiag, w.sc.p, w.partials.p, w.counter.p, mode, rtol2);
        else
            hipLaunchKernelGGL((cg_residual_kernel<V, PREC, NT, 8>), grid, blk, 0, s, n, A64.row.p, A64.col.p,
                               A64.val.p, b, x, rp, rs, cur, idiag, w.sc.p, w.partials.p, w.counter.p, mode, rtol2);
    };
    auto residual = [&](int mode, uint32_t it) {
        V2<V>* cur = vv.P[it & 1];
        if constexpr (PREC == kPrecLLt) {
            V2<V>* rqc = vv.RQ[it & 1];
            residual_kernel(reinterpret_cast<V*>(rqc), 2u, cur, mode);
            hipLaunchKernelGGL((llt_update_kernel<V, XV, CH, R, NT>), grid, blk, 0, s, MU, xacc, vv.D[it & 1], rqc, rqc,
                               vv.w, w.sc.p, w.partials.p, mode == 0 ? 1 : 2);
            hipLaunchKernelGGL((llt_lo_kernel<V, CH, R, NT>), grid, blk, 0, s, ML, vv.w, vv.D[(it + 1) & 1], vv.D[it & 1],
                               w.sc.p, w.partials.p, mode == 0 ? 1 : 2);
        } else {
            residual_kernel(vv.r, 1u, cur, mode);
        }
    };
    // one batch of o.batch iterations starting at iteration it0 (kernels
    // only: a captured batch contains no memset or copy nodes)
    auto launch_batch = [&](uint32_t it0) {
        for (uint32_t k = 0; k < o.batch; ++k) {
            const uint32_t it = it0 + k;
            V2<V>* cur = vv.P[it & 1];
            V2<V>* nxt = vv.P[(it + 1) & 1];
            if constexpr (PREC == kPrecLLt) {
                V2<V>* rqc = vv.RQ[it & 1];
                V2<V>* rqn = vv.RQ[(it + 1) & 1];
                V* dc = vv.D[it & 1];
                hipLaunchKernelGGL((llt_dir_kernel<V, DIFF, CH, R, NT>), grid, blk, 0, s, MA, dc,
                                   reinterpret_cast<V*>(rqc) + 1, 2u, w.sc.p, w.partials.p);
                hipLaunchKernelGGL((llt_update_kernel<V, XV, CH, R, NT>), grid, blk, 0, s, MU, xacc, dc, rqc, rqn, vv.w,
                                   w.sc.p, w.partials.p, 0);
                hipLaunchKernelGGL((llt_lo_kernel<V, CH, R, NT>), grid, blk, 0, s, ML, vv.w, dc, vv.D[(it + 1) & 1], w.sc.p,
                                   w.partials.p, 0);
            } else {
                hipLaunchKernelGGL((cg_dir_kernel<V, DIFF, CH, R, NT>), grid, blk, 0, s, MA, cur, nxt, vv.q, 1u, w.sc.p,
                                   w.partials.p);
                hipLaunchKernelGGL((cg_update_kernel<V, XV, PREC, NT>), grid, blk, 0, s, n, xacc, vv.r, nxt, vv.q, idiag,
                                   w.sc.p, w.partials.p, w.counter.p);
            }
        }
    };
    const bool use_graph = tuning().graph != 0;
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    residual(0, 0);
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
            residual(1, h.iter);
            ++reps;
            h = read_sc(w, s);
            rr_last_rep = h.aux;
            if (h.aux <= h.tol2 || h.iter >= o.maxit) break;
            continue;
        }
        const uint32_t it0 = h.iter;
        if (use_graph) {
            // the batch depends on it0 only through its parity (buffer
            // ping-pong), so one graph per parity serves the whole solve
            const uint32_t par = it0 & 1u;
            if (!exec[par]) {
                hipGraph_t gr = nullptr;
                SA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                launch_batch(par);
                SA_HIP(hipStreamEndCapture(s, &gr));
                SA_HIP(hipGraphInstantiate(&exec[par], gr, nullptr, nullptr, 0));
                SA_HIP(hipGraphDestroy(gr));
            }
            SA_HIP(hipGraphLaunch(exec[par], s));
        } else {
            launch_batch(it0);
        }
        h = read_sc(w, s);
        if (!h.done && o.replace_factor > 0 && h.rr < o.replace_factor * o.replace_factor * rr_last_rep) {
            residual(1, h.iter);
            ++reps;
            h = read_sc(w, s);
            rr_last_rep = h.aux;
        }
    }
    double tr[2];
    if (wide_rows)
        hipLaunchKernelGGL((true_residual_kernel<NT, 16>), grid, blk, 0, s, n, A64.row.p, A64.col.p, A64.val.p, b, x,
                           w.partials.p, w.counter.p, w.out2.p);
    else
        hipLaunchKernelGGL((true_residual_kernel<NT, 8>), grid, blk, 0, s, n, A64.row.p, A64.col.p, A64.val.p, b, x,
                           w.partials.p, w.counter.p, w.out2.p);
    d2h(tr, w.out2.p, 2);
    SA_HIP(hipEventRecord(e1, s));
    SA_HIP(hipEventSynchronize(e1));
    float ms = 0;
    SA_HIP(hipEventElapsedTime(&ms, e0, e1));
    SA_HIP(hipEventDestroy(e0));
    SA_HIP(hipEventDestroy(e1));
    for (auto& ge : exec)
        if (ge) SA_HIP(hipGraphExecDestroy(ge));
    h = read_sc(w, s);
    st.iters = h.iter;
    st.replacements = reps;
    st.true_rel = tr[1] > 0 ? std::sqrt(tr[0] / tr[1]) : std::sqrt(tr[0]);
    st.ms = ms;
}

}  // namespace

void true_residual(const DevCsr<double>& A, const double* b, const double* x, double* out2, CgWork& w,
                   hipStream_t s) {
    w.resize(A.n);
    if ((uint64_t)A.nnz > 8ull * A.n)
        hipLaunchKernelGGL((true_residual_kernel<kCgThreads, 16>), dim3(w.blocks), dim3(kCgThreads), 0, s, A.n,
                           A.row.p, A.col.p, A.val.p, b, x, w.partials.p, w.counter.p, w.out2.p);
    else
        hipLaunchKernelGGL((true_residual_kernel<kCgThreads, 8>), dim3(w.blocks), dim3(kCgThreads), 0, s, A.n,
                           A.row.p, A.col.p, A.val.p, b, x, w.partials.p, w.counter.p, w.out2.p);
    d2h(out2, w.out2.p, 2);
}

// SELL entries loaded per step of the row loop (all loads of a step are
// issued before their gathers); block size of the L L^T kernels.
template <class V, bool DIFF, int CH>
void cg_solve_llt(const DevCsr<double>& A64, const Sell<V>& S, const LowerFactor<V>* L, const double* b, double* x,
                  const CgOptions& o, CgWork& w, SolveStats& st, hipStream_t s) {
    const uint32_t nt = tuning().llt_threads;
    if (nt == 256)
        cg_drive<V, DIFF, kPrecLLt, CH, 1, 256>(A64, S, nullptr, L, b, x, o, w, st, s);
    else if (nt == 512)
        cg_drive<V, DIFF, kPrecLLt, CH, 1, 512>(A64, S, nullptr, L, b, x, o, w, st, s);
    else
        cg_drive<V, DIFF, kPrecLLt, CH, 1, 1024>(A64, S, nullptr, L, b, x, o, w, st, s);
}

template <class V, int CH>
void cg_solve_ch(const DevCsr<double>& A64, const Sell<V>& S, bool diff, int32_t prec, const V* idiag,
                 const LowerFactor<V>* L, const double* b, double* x, const CgOptions& o, CgWork& w,
                 SolveStats& st, hipStream_t s) {
    constexpr int NT = kCgThreads;
    if (diff) {
        if (prec == kPrecLLt)
            cg_solve_llt<V, true, CH>(A64, S, L, b, x, o, w, st, s);
        else if (prec == kPrecJacobi)
            cg_drive<V, true, kPrecJacobi, CH, 1, NT>(A64, S, idiag, L, b, x, o, w, st, s);
        else
            cg_drive<V, true, kPrecNone, CH, 1, NT>(A64, S, idiag, L, b, x, o, w, st, s);
    } else {
        if (prec == kPrecLLt)
            cg_solve_llt<V, false, CH>(A64, S, L, b, x, o, w, st, s);
        else if (prec == kPrecJacobi)
            cg_drive<V, false, kPrecJacobi, CH, 1, NT>(A64, S, idiag, L, b, x, o, w, st, s);
        else
            cg_drive<V, false, kPrecNone, CH, 1, NT>(A64, S, idiag, L, b, x, o, w, st, s);
    }
}

template <class V>
void cg_solve(const DevCsr<double>& A64, const Sell<V>& S, bool diff, int32_t prec, const V* idiag,
              const LowerFactor<V>* L, const double* b, double* x, const CgOptions& o, CgWork& w, SolveStats& st,
              hipStream_t s) {
    if (prec == kPrecLLt && !L) throw std::invalid_argument("L L^T preconditioner requested without a factor");
    if (prec == kPrecJacobi && !idiag) throw std::invalid_argument("Jacobi preconditioner requested without a diagonal");
    if (tuning().sell_chunk == 4)
        cg_solve_ch<V, 4>(A64, S, diff, prec, idiag, L, b, x, o, w, st, s);
    else
        cg_solve_ch<V, 8>(A64, S, diff, prec, idiag, L, b, x, o, w, st, s);
}
template void cg_solve<float>(const DevCsr<double>&, const Sell<float>&, bool, int32_t, const float*,
                              const LowerFactor<float>*, const double*, double*, const CgOptions&, CgWork&,
                              SolveStats&, hipStream_t);
template void cg_solve<double>(const DevCsr<double>&, const Sell<double>&, bool, int32_t, const double*,
                               const LowerFactor<double>*, const double*, double*, const CgOptions&, CgWork&,
                               SolveStats&, hipStream_t);

// One launch of a direction kernel on scratch vectors (ping-pong by k):
// the L L^T path's with NT-thread blocks, or the Jacobi path's.
template <class V, int CH, int NT>
void dir_launch(const SellArgs<V>& M, bool diff, bool llt, CgVecs<V>& vv, CgWork& w, uint32_t k, hipStream_t s) {
    const dim3 grid(w.blocks), blk(w.threads);
    V2<V>* cur = vv.P[k & 1];
    V2<V>* nxt = vv.P[(k + 1) & 1];
    V* q = reinterpret_cast<V*>(vv.RQ[k & 1]) + 1;
    if (llt && diff)
        hipLaunchKernelGGL((llt_dir_kernel<V, true, CH, 1, NT>), grid, blk, 0, s, M, vv.D[k & 1], q, 2u, w.sc.p,
                           w.partials.p);
    else if (llt)
        hipLaunchKernelGGL((llt_dir_kernel<V, false, CH, 1, NT>), grid, blk, 0, s, M, vv.D[k & 1], q, 2u, w.sc.p,
                           w.partials.p);
    else if (diff)
        hipLaunchKernelGGL((cg_dir_kernel<V, true, CH, 1, kCgThreads>), grid, blk, 0, s, M, cur, nxt, vv.q, 1u,
                           w.sc.p, w.partials.p);
    else
        hipLaunchKernelGGL((cg_dir_kernel<V, false, CH, 1, kCgThreads>), grid, blk, 0, s, M, cur, nxt, vv.q, 1u,
                           w.sc.p, w.partials.p);
}

// Average duration (ms, HIP events on the context stream) of a CG
// direction kernel over `reps` back-to-back launches on S: llt = false, the
// {z, d}-pair kernel of the Jacobi path (row-sum form when diff); llt = true,
// the L L^T path's kernel on a ready-made d.  The vectors are scratch (filled
// with ones).
template <class V>
double time_dir_kernel(const Sell<V>& S, bool diff, bool llt, uint32_t reps, CgWork& w, hipStream_t s) {
    const uint32_t n = S.n;
    w.resize(n, llt ? tuning().llt_threads : (uint32_t)kCgThreads);
    CgVecs<V> vv = carve<V>(w, n);
    CgScalars h{};
    h.maxit = 0xffffffffu;
    h.rz = 1.0;
    SA_HIP(hipMemcpyAsync(w.sc.p, &h, sizeof(h), hipMemcpyHostToDevice, s));
    const size_t nv = ((size_t)n + 63) & ~(size_t)63;
    fill<V>(reinterpret_cast<V*>(vv.P[0]), 10 * nv, (V)1, s);  // P, RQ and D
    const SellArgs<V> M = sell_args(S);
    const bool c4 = tuning().sell_chunk == 4;
    const uint32_t nt = w.threads;
    auto launch = [&](uint32_t k) {
        if (c4 && nt == 256)
            dir_launch<V, 4, 256>(M, diff, llt, vv, w, k, s);
        else if (c4 && nt == 512)
            dir_launch<V, 4, 512>(M, diff, llt, vv, w, k, s);
        else if (c4)
            dir_launch<V, 4, 1024>(M, diff, llt, vv, w, k, s);
        else if (nt == 256)
            dir_launch<V, 8, 256>(M, diff, llt, vv, w, k, s);
        else if (nt == 512)
            dir_launch<V, 8, 512>(M, diff, llt, vv, w, k, s);
        else
            dir_launch<V, 8, 1024>(M, diff, llt, vv, w, k, s);
    };
    for (uint32_t k = 0; k < 3; ++k) launch(k);
    hipEvent_t e0, e1;
    SA_HIP(hipEventCreate(&e0));
    SA_HIP(hipEventCreate(&e1));
    SA_HIP(hipEventRecord(e0, s));
    for (uint32_t k = 0; k < reps; ++k) launch(k);
    SA_HIP(hipEventRecord(e1, s));
    SA_HIP(hipEventSynchronize(e1));
    float ms = 0;
    SA_HIP(hipEventElapsedTime(&ms, e0, e1));
    SA_HIP(hipEventDestroy(e0));
    SA_HIP(hipEventDestroy(e1));
    return reps ? (double)ms / reps : 0.0;
}
template double time_dir_kernel<float>(const Sell<float>&, bool, bool, uint32_t, CgWork&, hipStream_t);
template double time_dir_kernel<double>(const Sell<double>&, bool, bool, uint32_t, CgWork&, hipStream_t);

// Average duration (ms) of the L L^T update kernel (x += alpha d,
// r' = r - alpha q, w = L^T r', partials of w.w and r'.r') or of the lower
// apply (z = L w, d_new = z + beta d_old) over `reps` launches on factor L
// with the configured block size.  The vectors are scratch (ones), the
// partials ones and the stop flag is never raised (maxit unbounded).
template <class V, int CH, int NT>
void update_launches(const LowerFactor<V>& L, uint32_t reps, CgVecs<V>& vv, double* xs, CgWork& w, hipStream_t s,
                     bool lower) {
    SellArgs<V> MU = sell_args(L.up), ML = sell_args(L.lo);
    MU.diag = L.diag.p;
    ML.diag = L.diag.p;
    for (uint32_t k = 0; k < reps; ++k) {
        if (lower)
            hipLaunchKernelGGL((llt_lo_kernel<V, CH, 1, NT>), dim3(w.blocks), dim3(NT), 0, s, ML, vv.w, vv.D[k & 1],
                               vv.D[(k + 1) & 1], w.sc.p, w.partials.p, 0);
        else
            hipLaunchKernelGGL((llt_update_kernel<V, double, CH, 1, NT>), dim3(w.blocks), dim3(NT), 0, s, MU, xs,
                               vv.D[k & 1], vv.RQ[k & 1], vv.RQ[(k + 1) & 1], vv.w, w.sc.p, w.partials.p, 0);
    }
}
template <class V>
double time_llt_update_kernel(const LowerFactor<V>& L, uint32_t reps, CgWork& w, hipStream_t s, bool lower) {
    const uint32_t n = L.n;
    w.resize(n, tuning().llt_threads);
    CgVecs<V> vv = carve<V>(w, n);
    CgScalars h{};
    h.maxit = 0xffffffffu;
    h.rz = 1.0;
    h.rz_old = 1.0;
    SA_HIP(hipMemcpyAsync(w.sc.p, &h, sizeof(h), hipMemcpyHostToDevice, s));
    const size_t nv = ((size_t)n + 63) & ~(size_t)63;
    fill<V>(reinterpret_cast<V*>(vv.P[0]), 10 * nv + 3 * nv, (V)1, s);  // P, RQ, D, r, q, w
    DBuf<double> xs;  // scratch solution
    xs.resize(n);
    fill<double>(xs.p, n, 0.0, s);
    fill<double>(w.partials.p, 3 * kMaxCgBlocks, 1.0, s);
    const bool c4 = tuning().sell_chunk == 4;
    const uint32_t nt = w.threads;
    auto launch = [&](uint32_t r) {
        if (c4 && nt == 256)
            update_launches<V, 4, 256>(L, r, vv, xs.p, w, s, lower);
        else if (c4 && nt == 512)
            update_launches<V, 4, 512>(L, r, vv, xs.p, w, s, lower);
        else if (c4)
            update_launches<V, 4, 1024>(L, r, vv, xs.p, w, s, lower);
        else if (nt == 256)
            update_launches<V, 8, 256>(L, r, vv, xs.p, w, s, lower);
        else if (nt == 512)
            update_launches<V, 8, 512>(L, r, vv, xs.p, w, s, lower);
        else
            update_launches<V, 8, 1024>(L, r, vv, xs.p, w, s, lower);
    };
    launch(3);
    hipEvent_t e0, e1;
    SA_HIP(hipEventCreate(&e0));
    SA_HIP(hipEventCreate(&e1));
    SA_HIP(hipEventRecord(e0, s));
    launch(reps);
    SA_HIP(hipEventRecord(e1, s));
    SA_HIP(hipEventSynchronize(e1));
    float ms = 0;
    SA_HIP(hipEventElapsedTime(&ms, e0, e1));
    SA_HIP(hipEventDestroy(e0));
    SA_HIP(hipEventDestroy(e1));
    return reps ? (double)ms / reps : 0.0;
}
template double time_llt_update_kernel<float>(const LowerFactor<float>&, uint32_t, CgWork&, hipStream_t, bool);

}  // namespace stream_amd
