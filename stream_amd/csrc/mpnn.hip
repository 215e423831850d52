// Learned L L^T factor: features, forward, Kaporin loss, backward, Adam and
// packing into the CG factor (see mpnn.hpp for the network).
#include <cmath>
#include <cstring>
#include <vector>

#include "mpnn.hpp"
#include "precond.hpp"

namespace stream_amd {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

constexpr float kDiagShift = 0.54132485461291810f;  // log(e - 1): softplus = 1
constexpr uint32_t kRsd = 3 * kHid;                  // row stride of [R|S|Dg]
constexpr uint32_t kPq = 2 * kHid;                   // row stride of [P|Q]
constexpr uint32_t kHeadPart = 448;                  // de_w 192, de_b 64, o_w 64, dd_b 64, d_w 64
constexpr uint32_t kMsgPart = 256;                   // e_w 192, m_b 64
constexpr uint32_t kEncPart = (kNodeFeats + 1) * kHid;  // enc_w 5 x 64, enc_b 64
constexpr uint32_t kEdgeGrid = 2048;                 // blocks of the partial-writing edge kernels
constexpr uint32_t kGemmGrid = 2048;
constexpr uint32_t kDwGrid = 512;

// BF16 [out][in] copies for the forward GEMMs.
constexpr uint32_t wt_pq(uint32_t l) { return l * 4u * kHid * kHid; }
constexpr uint32_t wt_u(uint32_t l) { return l * 4u * kHid * kHid + 2u * kHid * kHid; }
constexpr uint32_t kWtDec = kLayers * 4u * kHid * kHid;
constexpr uint32_t kWtTotal = kWtDec + 3u * kHid * kHid;

__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ float bfr(float x) { return bf2f(f2bf(x)); }

__device__ __forceinline__ float wsumf(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ bf16x8 pack8(const float4 a, const float4 b) {
    u16x8 u = {f2bf(a.x), f2bf(a.y), f2bf(a.z), f2bf(a.w), f2bf(b.x), f2bf(b.y), f2bf(b.z), f2bf(b.w)};
    return __builtin_bit_cast(bf16x8, u);
}

__device__ __forceinline__ uint32_t wave_id() { return threadIdx.x >> 6; }
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ float softplusf(float x) { return x > 20.f ? x : log1pf(expf(x)); }

// ---------------------------------------------------------------------------
// Features
// ---------------------------------------------------------------------------

// Per row: diagonal position, 1/sqrt(A_ii), degree; per block: box of the
// points and the sum of off-diagonal edge lengths (dim 2 or 3).
constexpr int kPrepVals = 8;  // xmin, xmax, ymin, ymax, zmin, zmax, length sum, count

__device__ __forceinline__ double edge_length(const double* __restrict__ xy, uint32_t xs, uint32_t dim, uint32_t lo,
                                              uint32_t hi) {
    const double dx = xy[(size_t)lo * xs] - xy[(size_t)hi * xs];
    const double dy = xy[(size_t)lo * xs + 1] - xy[(size_t)hi * xs + 1];
    double q = dx * dx + dy * dy;
    if (dim == 3) {
        const double dz = xy[(size_t)lo * xs + 2] - xy[(size_t)hi * xs + 2];
        q += dz * dz;
    }
    return sqrt(q);
}

__global__ __launch_bounds__(256) void prep_rows_kernel(uint32_t n, const uint32_t* __restrict__ row,
                                                        const uint32_t* __restrict__ col,
                                                        const double* __restrict__ val, const double* __restrict__ xy,
                                                        uint32_t xs, uint32_t dim, double* __restrict__ sd,
                                                        uint32_t* __restrict__ dpos, float* __restrict__ deg,
                                                        double* __restrict__ part) {
    __shared__ double red[kPrepVals][256];
    double bx[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
    double ls = 0.0, cnt = 0.0;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        for (uint32_t a = 0; a < dim; ++a) {
            const double c = xy[(size_t)i * xs + a];
            bx[2 * a] = fmin(bx[2 * a], c);
            bx[2 * a + 1] = fmax(bx[2 * a + 1], c);
        }
        uint32_t dp = ~0u, d = 0;
        double aii = 0.0;
        for (uint32_t e = row[i]; e < row[i + 1]; ++e) {
            const uint32_t j = col[e];
            if (j == i) {
                dp = e;
                aii = val[e];
                continue;
            }
            ++d;
            ls += edge_length(xy, xs, dim, i < j ? i : j, i < j ? j : i);
        }
        cnt += d;
        dpos[i] = dp;
        deg[i] = (float)d;
        sd[i] = aii > 0.0 ? 1.0 / sqrt(aii) : 1.0;
    }
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = bx[k];
    red[6][threadIdx.x] = ls;
    red[7][threadIdx.x] = cnt;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const uint32_t o = threadIdx.x + w;
            for (int k = 0; k < 6; k += 2) {
                red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][o]);
                red[k + 1][threadIdx.x] = fmax(red[k + 1][threadIdx.x], red[k + 1][o]);
            }
            red[6][threadIdx.x] += red[6][o];
            red[7][threadIdx.x] += red[7][o];
        }
        __syncthreads();
    }
    if (threadIdx.x < kPrepVals) part[(size_t)blockIdx.x * kPrepVals + threadIdx.x] = red[threadIdx.x][0];
}

// One block of kFoldThreads: thread t folds blocks t, t + kFoldThreads, ...
// in order, an xor butterfly combines each wave's lanes, and thread 0
// combines the waves in wave order (min/max exact, sums in a fixed order).
constexpr uint32_t kFoldThreads = 1024;
__global__ __launch_bounds__(kFoldThreads) void prep_fold_kernel(uint32_t nblk, const double* __restrict__ part,
                                                                 double* __restrict__ scal) {
    __shared__ double s_w[kFoldThreads / 64][kPrepVals];
    const uint32_t l = threadIdx.x;
    double b[kPrepVals] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY, 0.0, 0.0};
    for (uint32_t k = l; k < nblk; k += kFoldThreads) {
        const double* p = part + (size_t)k * kPrepVals;
        for (int c = 0; c < 6; c += 2) {
            b[c] = fmin(b[c], p[c]);
            b[c + 1] = fmax(b[c + 1], p[c + 1]);
        }
        b[6] += p[6];
        b[7] += p[7];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        for (int c = 0; c < 6; c += 2) {
            b[c] = fmin(b[c], __shfl_xor(b[c], o, 64));
            b[c + 1] = fmax(b[c + 1], __shfl_xor(b[c + 1], o, 64));
        }
        b[6] += __shfl_xor(b[6], o, 64);
        b[7] += __shfl_xor(b[7], o, 64);
    }
    if ((l & 63) == 0)
        for (int k = 0; k < kPrepVals; ++k) s_w[l / 64][k] = b[k];
    __syncthreads();
    if (l == 0) {
        for (uint32_t w = 1; w < kFoldThreads / 64; ++w) {
            for (int c = 0; c < 6; c += 2) {
                b[c] = fmin(b[c], s_w[w][c]);
                b[c + 1] = fmax(b[c + 1], s_w[w][c + 1]);
            }
            b[6] += s_w[w][6];
            b[7] += s_w[w][7];
        }
        for (int k = 0; k < kPrepVals; ++k) scal[k] = b[k];
    }
}

// Node features, edge features (bf16-rounded) and the transpose map.
__global__ __launch_bounds__(256) void prep_feat_kernel(uint32_t n, uint32_t lanes, const uint32_t* __restrict__ row,
                                                        const uint32_t* __restrict__ col,
                                                        const double* __restrict__ val, const double* __restrict__ xy,
                                                        uint32_t xs, uint32_t dim, const uint8_t* __restrict__ bnd,
                                                        const double* __restrict__ sd,
                                                        const uint32_t* __restrict__ dpos,
                                                        const double* __restrict__ scal, float* __restrict__ nfeat,
                                                        float2* __restrict__ efeat, uint32_t* __restrict__ tr,
                                                        uint32_t* __restrict__ flags) {
    double ext = fmax(scal[1] - scal[0], scal[3] - scal[2]);
    if (dim == 3) ext = fmax(ext, scal[5] - scal[4]);
    if (!(ext > 0.0)) ext = 1.0;
    const double lbar = scal[7] > 0.0 ? scal[6] / scal[7] : 1.0;
    // `lanes` lanes (8 or 16) per row: the row's entries in consecutive
    // lanes, so the edge-feature stores of a wavefront are whole segments
    const uint32_t sub = threadIdx.x % lanes;
    const uint32_t rows_per_block = 256 / lanes;
    for (uint32_t i = blockIdx.x * rows_per_block + threadIdx.x / lanes; i < n; i += gridDim.x * rows_per_block) {
        if (sub == 0) {
            const double aii = dpos[i] != ~0u ? val[dpos[i]] : 0.0;
            float* nf = nfeat + (size_t)i * kNodeFeats;
            for (uint32_t a = 0; a < 3; ++a)
                nf[a] = a < dim ? (float)((xy[(size_t)i * xs + a] - scal[2 * a]) / ext) : 0.f;
            nf[3] = bnd[i] ? 1.f : 0.f;
            nf[4] = (float)(aii / 8.0);
        }
        for (uint32_t e = row[i] + sub; e < row[i + 1]; e += lanes) {
            const uint32_t j = col[e];
            if (j == i) {
                efeat[e] = make_float2(0.f, 0.f);
                if (tr) tr[e] = e;
                continue;
            }
            const uint32_t lo = i < j ? i : j, hi = i < j ? j : i;
            const double ah = val[e] * (sd[lo] * sd[hi]);
            const double len = edge_length(xy, xs, dim, lo, hi) / lbar;
            efeat[e] = make_float2(bfr((float)ah), bfr((float)len));
            if (!tr) continue;  // inference: no transpose map
            // (j, i) in row j: binary search (columns sorted)
            uint32_t a = row[j], b = row[j + 1];
            while (a < b) {
                const uint32_t m = (a + b) >> 1;
                if (col[m] < i)
                    a = m + 1;
                else
                    b = m;
            }
            if (a == row[j + 1] || col[a] != i) {
                raise_flag(flags, kErrArgument);  // pattern not symmetric
                a = e;
            }
            tr[e] = a;
        }
    }
}

// ---------------------------------------------------------------------------
// Dense layers on MFMA
// ---------------------------------------------------------------------------

enum : int { kEpiStore = 0, kEpiResRelu = 1, kEpiAdd = 2 };

struct GemmArgs {
    uint32_t n;
    const float* x0;
    uint32_t ld0;
    const float* x1;
    uint32_t ld1;
    const uint16_t* wt;  // [N][K] bf16
    const float* bias;   // [N] or null
    float* y;
    uint32_t ldy;
    const float* add0;
    uint32_t lda0;
    const float* add1;
    uint32_t lda1;
    float* u;  // kEpiResRelu: pre-activation [n][N] (null: not stored)
};

// Y[n][N] = bf16(X[n][K]) * bf16(W) with FP32 accumulation; X columns
// [0, K0) come from x0, [K0, K) from x1.  Four waves per block, 16 rows per
// wave, the whole weight matrix in LDS; v_mfma_f32_16x16x32_bf16 tiles:
// lane l holds A[l%16][8(l/16)..+7], B[8(l/16)..+7][l%16] and D rows
// 4(l/16)..+3 of column l%16.
template <int K, int K0, int N, int EPI>
__global__ __launch_bounds__(256) void gemm_rows_kernel(GemmArgs a) {
    constexpr int KP = K + 8;
    __shared__ __attribute__((aligned(16))) uint16_t sw[N * KP];
    for (int t = threadIdx.x; t < N * (K / 8); t += 256) {
        const int r = t / (K / 8), c = (t % (K / 8)) * 8;
        *reinterpret_cast<uint4*>(&sw[r * KP + c]) = *reinterpret_cast<const uint4*>(&a.wt[r * K + c]);
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lr = lane & 15, lk = lane >> 4;
    for (uint32_t tile = blockIdx.x; (size_t)tile * 64 < a.n; tile += gridDim.x) {
        const uint32_t r0 = tile * 64 + wave * 16;
        const uint32_t r = r0 + lr;
        bf16x8 av[K / 32];
#pragma unroll
        for (int kk = 0; kk < K / 32; ++kk) {
            const int k = kk * 32 + lk * 8;
            float4 v0 = make_float4(0.f, 0.f, 0.f, 0.f), v1 = v0;
            if (r < a.n) {
                const float* src = k < K0 ? a.x0 + (size_t)r * a.ld0 + k : a.x1 + (size_t)r * a.ld1 + (k - K0);
                v0 = *reinterpret_cast<const float4*>(src);
                v1 = *reinterpret_cast<const float4*>(src + 4);
            }
            av[kk] = pack8(v0, v1);
        }
#pragma unroll
        for (int nt = 0; nt < N / 16; ++nt) {
            floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < K / 32; ++kk) {
                const bf16x8 bv = *reinterpret_cast<const bf16x8*>(&sw[(nt * 16 + lr) * KP + kk * 32 + lk * 8]);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[kk], bv, acc, 0, 0, 0);
            }
            const int c = nt * 16 + lr;
            const float bc = a.bias ? a.bias[c] : 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t rr = r0 + lk * 4 + i;
                if (rr >= a.n) continue;
                float v = acc[i] + bc;
                if (EPI == kEpiResRelu) {
                    if (a.u) a.u[(size_t)rr * N + c] = v;
                    v = a.add0[(size_t)rr * a.lda0 + c] + fmaxf(v, 0.f);
                } else if (EPI == kEpiAdd) {
                    if (a.add0) v += a.add0[(size_t)rr * a.lda0 + c];
                    if (a.add1) v += a.add1[(size_t)rr * a.lda1 + c];
                }
                a.y[(size_t)rr * a.ldy + c] = v;
            }
        }
    }
}

// Weight gradient partials: partial[block] = bf16(X)^T bf16(dY) over the
// block's rows ([K][N], then the column sums of dY when BIAS).  Rows are
// staged 32 at a time into LDS transposed, so each MFMA operand is one
// 16-byte LDS read per lane.
template <int K, int K0, int N, bool BIAS>
__global__ __launch_bounds__(256) void gemm_dw_kernel(uint32_t n, uint32_t rows_per_block, const float* __restrict__ x0,
                                                      uint32_t ld0, const float* __restrict__ x1, uint32_t ld1,
                                                      const float* __restrict__ dy, uint32_t lddy,
                                                      float* __restrict__ partial) {
    constexpr int RC = 32, RP = RC + 8;
    constexpr int NT = N / 16, TILES = (K / 16) * NT, TPW = TILES / 4;
    static_assert(TILES % 4 == 0, "tiles split over four waves");
    static_assert(!BIAS || N == 64, "bias sums assume 64 columns");
    __shared__ __attribute__((aligned(16))) uint16_t sx[K * RP];
    __shared__ __attribute__((aligned(16))) uint16_t sy[N * RP];
    __shared__ float sb[256];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    floatx4 acc[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;
    const uint32_t rb = blockIdx.x * rows_per_block;
    const uint32_t re = rb + rows_per_block < n ? rb + rows_per_block : n;
    for (uint32_t c0 = rb; c0 < re; c0 += RC) {
        __syncthreads();
        for (int e = threadIdx.x; e < RC * K; e += 256) {
            const int r = e / K, k = e % K;
            const uint32_t rw = c0 + r;
            float v = 0.f;
            if (rw < re) v = k < K0 ? x0[(size_t)rw * ld0 + k] : x1[(size_t)rw * ld1 + (k - K0)];
            sx[k * RP + r] = f2bf(v);
        }
        for (int e = threadIdx.x; e < RC * N; e += 256) {
            const int r = e / N, c = e % N;
            const uint32_t rw = c0 + r;
            const float v = rw < re ? dy[(size_t)rw * lddy + c] : 0.f;
            if (BIAS) bsum += v;
            sy[c * RP + r] = f2bf(v);
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
            const int tile = wave * TPW + t, kt = tile / NT, nt = tile % NT;
            const bf16x8 av = *reinterpret_cast<const bf16x8*>(&sx[(kt * 16 + lr) * RP + lk * 8]);
            const bf16x8 bv = *reinterpret_cast<const bf16x8*>(&sy[(nt * 16 + lr) * RP + lk * 8]);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[t], 0, 0, 0);
        }
    }
    float* out = partial + (size_t)blockIdx.x * (K * N + (BIAS ? N : 0));
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const int tile = wave * TPW + t, kt = tile / NT, nt = tile % NT;
#pragma unroll
        for (int i = 0; i < 4; ++i) out[(kt * 16 + lk * 4 + i) * N + nt * 16 + lr] = acc[t][i];
    }
    if (BIAS) {
        sb[threadIdx.x] = bsum;
        __syncthreads();
        if (threadIdx.x < N) {
            float s = 0.f;
            for (int q = 0; q < 256 / N; ++q) s += sb[threadIdx.x + q * N];
            out[K * N + threadIdx.x] = s;
        }
    }
}

__global__ __launch_bounds__(256) void reduce_partials_kernel(const float* __restrict__ partial, uint32_t nblk,
                                                              uint32_t stride, uint32_t width,
                                                              float* __restrict__ dst) {
    const uint32_t o = blockIdx.x * 256 + threadIdx.x;
    if (o >= width) return;
    double s = 0.0;
    for (uint32_t b = 0; b < nblk; ++b) s += partial[(size_t)b * stride + o];
    dst[o] = (float)s;
}

// ---------------------------------------------------------------------------
// Encoder (VALU: kNodeFeats inputs)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void enc_fwd_kernel(uint32_t n, const float* __restrict__ nfeat,
                                                      const uint16_t* __restrict__ wb, const float* __restrict__ w,
                                                      float* __restrict__ h0) {
    const uint32_t f = threadIdx.x & 63;
    const float b = w[kParams.enc_b + f];
    float wc[kNodeFeats];
#pragma unroll
    for (int c = 0; c < (int)kNodeFeats; ++c) wc[c] = bf2f(wb[kParams.enc_w + c * kHid + f]);
    for (uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6); i < n; i += gridDim.x * 4) {
        float z = 0.f;
#pragma unroll
        for (int c = 0; c < (int)kNodeFeats; ++c) z += bfr(nfeat[(size_t)i * kNodeFeats + c]) * wc[c];
        h0[(size_t)i * kHid + f] = fmaxf(z + b, 0.f);
    }
}

__global__ __launch_bounds__(256) void enc_bwd_kernel(uint32_t n, const float* __restrict__ nfeat,
                                                      const float* __restrict__ h0, const float* __restrict__ dh,
                                                      float* __restrict__ partial) {
    constexpr int K = kNodeFeats + 1;
    __shared__ float red[K][256];
    const uint32_t f = threadIdx.x & 63;
    float a[K];
#pragma unroll
    for (int c = 0; c < K; ++c) a[c] = 0.f;
    for (uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6); i < n; i += gridDim.x * 4) {
        const float d = h0[(size_t)i * kHid + f] > 0.f ? dh[(size_t)i * kHid + f] : 0.f;
#pragma unroll
        for (int c = 0; c < (int)kNodeFeats; ++c) a[c] += bfr(nfeat[(size_t)i * kNodeFeats + c]) * d;
        a[kNodeFeats] += d;
    }
#pragma unroll
    for (int c = 0; c < K; ++c) red[c][threadIdx.x] = a[c];
    __syncthreads();
    if (threadIdx.x < 64) {
        float* out = partial + (size_t)blockIdx.x * kEncPart;
#pragma unroll
        for (int c = 0; c < K; ++c)
            out[c * kHid + f] = red[c][f] + red[c][f + 64] + red[c][f + 128] + red[c][f + 192];
    }
}

// ---------------------------------------------------------------------------
// Messages (one wave per node, lane = feature)
// ---------------------------------------------------------------------------
constexpr uint32_t kMsgBatch = 8;  // edges whose loads are issued together

__global__ __launch_bounds__(256) void msg_fwd_kernel(uint32_t n, const uint32_t* __restrict__ row,
                                                      const uint32_t* __restrict__ col,
                                                      const float2* __restrict__ efeat,
                                                      const float* __restrict__ deg, const float* __restrict__ pq,
                                                      const uint16_t* __restrict__ wb, const float* __restrict__ w,
                                                      uint32_t l, float* __restrict__ m) {
    const uint32_t f = threadIdx.x & 63;
    const float v0 = bf2f(wb[kParams.e_w[l] + f]), v1 = bf2f(wb[kParams.e_w[l] + kHid + f]),
                v2 = bf2f(wb[kParams.e_w[l] + 2 * kHid + f]);
    const float b = w[kParams.m_b[l] + f];
    for (uint32_t i0 = blockIdx.x * 4 + wave_id(); i0 < n; i0 += gridDim.x * 4) {
        const uint32_t i = uniform(i0);
        const float p = pq[(size_t)i * kPq + f];
        float acc = 0.f;
        const uint32_t e1 = row[i + 1];
        // the row's columns, edge features and neighbour rows of up to kMsgBatch
        // edges are loaded before any is used; the adds keep the edge order
        for (uint32_t eb = row[i]; eb < e1; eb += kMsgBatch) {
            uint32_t jj[kMsgBatch];
            float2 ef[kMsgBatch];
            float qj[kMsgBatch];
#pragma unroll
            for (uint32_t u = 0; u < kMsgBatch; ++u) {
                const bool ok = eb + u < e1;
                jj[u] = ok ? col[eb + u] : i;
                ef[u] = ok ? efeat[eb + u] : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (uint32_t u = 0; u < kMsgBatch; ++u) qj[u] = jj[u] != i ? pq[(size_t)jj[u] * kPq + kHid + f] : 0.f;
#pragma unroll
            for (uint32_t u = 0; u < kMsgBatch; ++u) {
                const uint32_t j = jj[u];
                if (j == i) continue;  // the diagonal and the padding
                const float et = ef[u].x * v0 + ef[u].y * v1 + (j > i ? v2 : -v2);
                const float z = p + qj[u] + et + b;
                acc += fmaxf(z, 0.f);
            }
        }
        const float d = deg[i];
        m[(size_t)i * kHid + f] = d > 0.f ? acc / d : 0.f;
    }
}

// dP_j = sum over messages j <- i, dQ_j = sum over messages i <- j of
// dz = (dM_receiver / deg_receiver) [z > 0]; weight partials from the dP side.
__global__ __launch_bounds__(256) void msg_bwd_kernel(uint32_t n, const uint32_t* __restrict__ row,
                                                      const uint32_t* __restrict__ col,
                                                      const float2* __restrict__ efeat,
                                                      const float* __restrict__ deg, const float* __restrict__ pq,
                                                      const float* __restrict__ dx, const uint16_t* __restrict__ wb,
                                                      const float* __restrict__ w, uint32_t l,
                                                      float* __restrict__ dpq, float* __restrict__ partial) {
    __shared__ float red[4][256];
    const uint32_t f = threadIdx.x & 63;
    const float v0 = bf2f(wb[kParams.e_w[l] + f]), v1 = bf2f(wb[kParams.e_w[l] + kHid + f]),
                v2 = bf2f(wb[kParams.e_w[l] + 2 * kHid + f]);
    const float b = w[kParams.m_b[l] + f];
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, ab = 0.f;
    for (uint32_t j0 = blockIdx.x * 4 + wave_id(); j0 < n; j0 += gridDim.x * 4) {
        const uint32_t j = uniform(j0);
        const float pj = pq[(size_t)j * kPq + f], qj = pq[(size_t)j * kPq + kHid + f];
        const float dj = deg[j];
        const float gj = dj > 0.f ? dx[(size_t)j * kPq + kHid + f] / dj : 0.f;
        float dp = 0.f, dq = 0.f;
        for (uint32_t e = row[j]; e < row[j + 1]; ++e) {
            const uint32_t i = col[e];
            if (i == j) continue;
            const float2 ef = efeat[e];
            const float s2 = i > j ? v2 : -v2;
            // message j <- i
            const float z1 = pj + pq[(size_t)i * kPq + kHid + f] + (ef.x * v0 + ef.y * v1 + s2) + b;
            const float dz1 = z1 > 0.f ? gj : 0.f;
            dp += dz1;
            a0 += ef.x * dz1;
            a1 += ef.y * dz1;
            a2 += (i > j ? dz1 : -dz1);
            ab += dz1;
            // message i <- j
            const float z2 = pq[(size_t)i * kPq + f] + qj + (ef.x * v0 + ef.y * v1 - s2) + b;
            const float di = deg[i];
            const float gi = dx[(size_t)i * kPq + kHid + f] / di;
            dq += z2 > 0.f ? gi : 0.f;
        }
        dpq[(size_t)j * kPq + f] = dp;
        dpq[(size_t)j * kPq + kHid + f] = dq;
    }
    red[0][threadIdx.x] = a0;
    red[1][threadIdx.x] = a1;
    red[2][threadIdx.x] = a2;
    red[3][threadIdx.x] = ab;
    __syncthreads();
    if (threadIdx.x < 64) {
        float* out = partial + (size_t)blockIdx.x * kMsgPart;
#pragma unroll
        for (int c = 0; c < 4; ++c)
            out[c * kHid + f] = red[c][f] + red[c][f + 64] + red[c][f + 128] + red[c][f + 192];
    }
}

__global__ __launch_bounds__(256) void relu_mask_kernel(size_t total, const float* __restrict__ dh,
                                                        const float* __restrict__ u, float* __restrict__ du) {
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (size_t)gridDim.x * 256)
        du[t] = u[t] > 0.f ? dh[t] : 0.f;
}

// ---------------------------------------------------------------------------
// Output head
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void head_fwd_kernel(uint32_t n, const uint32_t* __restrict__ row,
                                                       const uint32_t* __restrict__ col,
                                                       const double* __restrict__ val,
                                                       const uint32_t* __restrict__ dpos,
                                                       const uint8_t* __restrict__ bnd,
                                                       const float2* __restrict__ efeat,
                                                       const float* __restrict__ rsd, const uint16_t* __restrict__ wb,
                                                       const float* __restrict__ w, float* __restrict__ gval,
                                                       float* __restrict__ od) {
    const uint32_t f = threadIdx.x & 63;
    const float v0 = bf2f(wb[kParams.de_w + f]), v1 = bf2f(wb[kParams.de_w + kHid + f]),
                v2 = bf2f(wb[kParams.de_w + 2 * kHid + f]);
    const float bde = w[kParams.de_b + f], wo = bf2f(wb[kParams.o_w + f]);
    const float bdd = w[kParams.dd_b + f], wd = bf2f(wb[kParams.d_w + f]);
    for (uint32_t k0 = blockIdx.x * 4 + wave_id(); k0 < n; k0 += gridDim.x * 4) {
        const uint32_t k = uniform(k0);
        const float r = rsd[(size_t)k * kRsd + f];
        const uint32_t e1 = row[k + 1];
        // columns are sorted, so the upper entries (the only ones with an
        // output) follow the diagonal
        const uint32_t dk = dpos[k];
        const uint32_t e0 = dk != ~0u ? dk + 1 : row[k];
        // loads of up to kMsgBatch edges are issued before any is used
        for (uint32_t eb = e0; eb < e1; eb += kMsgBatch) {
            uint32_t jj[kMsgBatch];
            float2 ef[kMsgBatch];
            float sj[kMsgBatch];
            bool nz[kMsgBatch];
#pragma unroll
            for (uint32_t u = 0; u < kMsgBatch; ++u) {
                const bool ok = eb + u < e1;
                jj[u] = ok ? col[eb + u] : k;
                ef[u] = ok ? efeat[eb + u] : make_float2(0.f, 0.f);
                nz[u] = ok && val[eb + u] != 0.0;
            }
            // only upper neighbours (j > k) carry an output: the lower part,
            // the diagonal and the padding load nothing
#pragma unroll
            for (uint32_t u = 0; u < kMsgBatch; ++u) sj[u] = jj[u] > k ? rsd[(size_t)jj[u] * kRsd + kHid + f] : 0.f;
#pragma unroll
            for (uint32_t u = 0; u < kMsgBatch; ++u) {
                if (jj[u] <= k) continue;  // lower part, diagonal and padding
                const float z = r + sj[u] + (ef[u].x * v0 + ef[u].y * v1 + v2) + bde;
                const float o = wsumf(bfr(fmaxf(z, 0.f)) * wo);
                if (f == 0) gval[eb + u] = nz[u] ? o : 0.f;
            }
        }
        const float zd = rsd[(size_t)k * kRsd + 2 * kHid + f] + bdd;
        const float o = wsumf(bfr(fmaxf(zd, 0.f)) * wd) + kDiagShift;
        if (f == 0) {
            od[k] = o;
            if (dk != ~0u) gval[dk] = bnd[k] ? 1.f : softplusf(o);
        }
    }
}

// Loss: y = (Ghat Ahat) at the pattern of row k of Ghat, partial T and
// sum log Ghat_kk.  One thread per row (FP64).
__global__ __launch_bounds__(256) void loss_kernel(uint32_t n, const uint32_t* __restrict__ row,
                                                   const uint32_t* __restrict__ col, const double* __restrict__ val,
                                                   const double* __restrict__ sd, const uint32_t* __restrict__ dpos,
                                                   const float* __restrict__ gval, double* __restrict__ y,
                                                   double* __restrict__ part, uint32_t* __restrict__ flags) {
    __shared__ double red[2][256];
    double tp = 0.0, lp = 0.0;
    for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
        if (dpos[k] == ~0u) {
            raise_flag(flags, kErrArgument);
            continue;
        }
        uint32_t P[kMaxUpper + 1], E[kMaxUpper + 1];
        double gP[kMaxUpper + 1];
        uint32_t m = 1;
        P[0] = k;
        E[0] = dpos[k];
        gP[0] = gval[dpos[k]];
        for (uint32_t e = row[k]; e < row[k + 1]; ++e) {
            const uint32_t c = col[e];
            if (c <= k) continue;
            if (m > kMaxUpper) {
                raise_flag(flags, kErrCapacity);
                break;
            }
            P[m] = c;
            E[m] = e;
            gP[m] = gval[e];
            ++m;
        }
        for (uint32_t a = 0; a < m; ++a) {
            const uint32_t pa = P[a];
            double ya = 0.0;
            for (uint32_t e = row[pa]; e < row[pa + 1]; ++e) {
                const uint32_t c = col[e];
                for (uint32_t b = 0; b < m; ++b)
                    if (P[b] == c) {
                        const uint32_t lo = pa < c ? pa : c, hi = pa < c ? c : pa;
                        ya += val[e] * (sd[lo] * sd[hi]) * gP[b];
                    }
            }
            y[E[a]] = ya;
            tp += gP[a] * ya;
        }
        lp += log(gP[0]);
    }
    red[0][threadIdx.x] = tp;
    red[1][threadIdx.x] = lp;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[(size_t)blockIdx.x * 2] = red[0][0];
        part[(size_t)blockIdx.x * 2 + 1] = red[1][0];
    }
}

// One wave: lane q sums blocks q, q + 64, ... in order, then an xor
// butterfly combines the lanes (a fixed order: deterministic).
__global__ __launch_bounds__(64) void loss_fold_kernel(uint32_t nblk, uint32_t n, const double* __restrict__ part,
                                                       double* __restrict__ lsc) {
    double t = 0.0, l = 0.0;
    for (uint32_t b = threadIdx.x; b < nblk; b += 64) {
        t += part[(size_t)b * 2];
        l += part[(size_t)b * 2 + 1];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        t += __shfl_xor(t, o, 64);
        l += __shfl_xor(l, o, 64);
    }
    if (threadIdx.x != 0) return;
    lsc[0] = t;
    lsc[1] = l;
    lsc[2] = log(t / (double)n) - 2.0 / (double)n * l;
    lsc[3] = (double)n;
}

// Head backward: dR (own upper edges), dS (edges (j, k) with j < k, via the
// transpose map), dDg; weight partials in the order de_w, de_b, o_w, dd_b, d_w.
__global__ __launch_bounds__(256) void head_bwd_kernel(uint32_t n, const uint32_t* __restrict__ row,
                                                       const uint32_t* __restrict__ col,
                                                       const double* __restrict__ val,
                                                       const uint32_t* __restrict__ dpos,
                                                       const uint32_t* __restrict__ tr,
                                                       const uint8_t* __restrict__ bnd,
                                                       const float2* __restrict__ efeat,
                                                       const float* __restrict__ rsd, const uint16_t* __restrict__ wb,
                                                       const float* __restrict__ w, const float* __restrict__ gval,
                                                       const float* __restrict__ od, const double* __restrict__ y,
                                                       const double* __restrict__ lsc, float* __restrict__ drsd,
                                                       float* __restrict__ partial) {
    __shared__ float red[7][256];
    const uint32_t f = threadIdx.x & 63;
    const float v0 = bf2f(wb[kParams.de_w + f]), v1 = bf2f(wb[kParams.de_w + kHid + f]),
                v2 = bf2f(wb[kParams.de_w + 2 * kHid + f]);
    const float bde = w[kParams.de_b + f], wo = bf2f(wb[kParams.o_w + f]);
    const float bdd = w[kParams.dd_b + f], wd = bf2f(wb[kParams.d_w + f]);
    const double cT = 2.0 / lsc[0], cn = 2.0 / (double)n;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, ab = 0.f, awo = 0.f, abd = 0.f, awd = 0.f;
    for (uint32_t k0 = blockIdx.x * 4 + wave_id(); k0 < n; k0 += gridDim.x * 4) {
        const uint32_t k = uniform(k0);
        const float r = rsd[(size_t)k * kRsd + f], s = rsd[(size_t)k * kRsd + kHid + f];
        float dr = 0.f, ds = 0.f;
        for (uint32_t e = row[k]; e < row[k + 1]; ++e) {
            const uint32_t j = col[e];
            if (j == k || val[e] == 0.0) continue;
            const float2 ef = efeat[e];
            const float et = ef.x * v0 + ef.y * v1 + v2;
            if (j > k) {
                const float gk = (float)(cT * y[e]);
                const float z = r + rsd[(size_t)j * kRsd + kHid + f] + et + bde;
                const float dz = z > 0.f ? gk * wo : 0.f;
                dr += dz;
                awo += gk * bfr(fmaxf(z, 0.f));
                a0 += ef.x * dz;
                a1 += ef.y * dz;
                a2 += dz;
                ab += dz;
            } else {
                const float gk = (float)(cT * y[tr[e]]);
                const float z = rsd[(size_t)j * kRsd + f] + s + et + bde;
                ds += z > 0.f ? gk * wo : 0.f;
            }
        }
        float dzd = 0.f;
        const uint32_t dp = dpos[k];
        if (!bnd[k] && dp != ~0u) {
            const float gd = (float)(cT * y[dp] - cn / (double)gval[dp]);
            const float o = od[k];
            const float dod = gd * (1.f / (1.f + expf(-o)));
            const float zd = rsd[(size_t)k * kRsd + 2 * kHid + f] + bdd;
            dzd = zd > 0.f ? dod * wd : 0.f;
            awd += dod * bfr(fmaxf(zd, 0.f));
            abd += dzd;
        }
        drsd[(size_t)k * kRsd + f] = dr;
        drsd[(size_t)k * kRsd + kHid + f] = ds;
        drsd[(size_t)k * kRsd + 2 * kHid + f] = dzd;
    }
    red[0][threadIdx.x] = a0;
    red[1][threadIdx.x] = a1;
    red[2][threadIdx.x] = a2;
    red[3][threadIdx.x] = ab;
    red[4][threadIdx.x] = awo;
    red[5][threadIdx.x] = abd;
    red[6][threadIdx.x] = awd;
    __syncthreads();
    if (threadIdx.x < 64) {
        float* out = partial + (size_t)blockIdx.x * kHeadPart;
#pragma unroll
        for (int c = 0; c < 7; ++c)
            out[c * kHid + f] = red[c][f] + red[c][f + 64] + red[c][f + 128] + red[c][f + 192];
    }
}

// ---------------------------------------------------------------------------
// Optimiser and weight copies
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void adam_kernel(uint32_t np, float* __restrict__ w, float* __restrict__ m,
                                                   float* __restrict__ v, const float* __restrict__ g, float lr,
                                                   float b1, float b2, float eps, float c1, float c2,
                                                   uint16_t* __restrict__ wb) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= np) return;
    const float gp = g[p];
    const float mp = b1 * m[p] + (1.f - b1) * gp;
    const float vp = b2 * v[p] + (1.f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    const float wp = w[p] - lr * (mp / c1) / (sqrtf(vp / c2) + eps);
    w[p] = wp;
    wb[p] = f2bf(wp);
}

__global__ __launch_bounds__(256) void to_bf16_kernel(uint32_t np, const float* __restrict__ w,
                                                      uint16_t* __restrict__ wb) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p < np) wb[p] = f2bf(w[p]);
}

// [in][out] -> [out][in] copies of pq_w, u_w, dec_w.
__global__ __launch_bounds__(256) void transpose_weights_kernel(const uint16_t* __restrict__ wb,
                                                                uint16_t* __restrict__ wt) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= kWtTotal) return;
    uint32_t src, dst, rows, cols;  // source [rows][cols]
    if (t >= kWtDec) {
        src = kParams.dec_w;
        dst = kWtDec;
        rows = kHid;
        cols = 3 * kHid;
    } else {
        const uint32_t l = t / (4 * kHid * kHid);
        if (t - wt_pq(l) < 2 * kHid * kHid) {
            src = kParams.pq_w[l];
            dst = wt_pq(l);
            rows = kHid;
            cols = 2 * kHid;
        } else {
            src = kParams.u_w[l];
            dst = wt_u(l);
            rows = 2 * kHid;
            cols = kHid;
        }
    }
    const uint32_t o = t - dst;      // position in [cols][rows]
    const uint32_t c = o / rows, r = o % rows;
    wt[t] = wb[src + r * cols + c];
}

// L = D^-1/2 Ghat^T into the SELL factor: the up slot of CSR entry (k, j),
// j > k, and the lo slot of entry (j, k) hold L_jk = sd_j Ghat_kj; diag
// L_kk = sd_k Ghat_kk.
__global__ __launch_bounds__(256) void pack_kernel(uint32_t n, const uint32_t* __restrict__ row,
                                                   const uint32_t* __restrict__ col, const double* __restrict__ sd,
                                                   const uint32_t* __restrict__ dpos, const float* __restrict__ gval,
                                                   const uint32_t* __restrict__ slot, float* __restrict__ up_val,
                                                   float* __restrict__ lo_val, float* __restrict__ diag,
                                                   uint32_t* __restrict__ flags) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    bool bad = false;
    const uint32_t dp = dpos[k];
    const float gd = dp != ~0u ? gval[dp] : 1.f;
    const float lkk = (float)(sd[k] * (double)gd);
    bad |= !(lkk > 0.f) || !isfinite(lkk);
    diag[k] = lkk;
    for (uint32_t e = row[k]; e < row[k + 1]; ++e) {
        const uint32_t j = col[e];
        if (j <= k) continue;
        const float v = (float)(sd[j] * (double)gval[e]);
        bad |= !isfinite(v);
        up_val[slot[e]] = v;
        const uint32_t f = csr_find(row, col, j, k);  // (j, k): the lo slot of L_jk
        if (f != ~0u) lo_val[slot[f]] = v;
    }
    if (bad) atomicMax(&flags[kFlagLearned], 1u);
}

unsigned blocks_for(size_t items, unsigned per_block, unsigned cap) {
    size_t b = (items + per_block - 1) / per_block;
    if (b < 1) b = 1;
    return (unsigned)(b < cap ? b : cap);
}

template <int K, int K0, int N, int EPI> void run_gemm(const GemmArgs& a, hipStream_t s) {
    if (!a.n) return;
    hipLaunchKernelGGL((gemm_rows_kernel<K, K0, N, EPI>), dim3(blocks_for(a.n, 64, kGemmGrid)), dim3(256), 0, s, a);
}

GemmArgs gargs(uint32_t n, const float* x0, uint32_t ld0, const uint16_t* wt, float* y, uint32_t ldy) {
    GemmArgs a{};
    a.n = n;
    a.x0 = x0;
    a.ld0 = ld0;
    a.wt = wt;
    a.y = y;
    a.ldy = ldy;
    return a;
}

template <int K, int K0, int N, bool BIAS>
void run_dw(uint32_t n, const float* x0, uint32_t ld0, const float* x1, uint32_t ld1, const float* dy,
            uint32_t lddy, float* partial, float* dst, hipStream_t s) {
    const uint32_t nb = blocks_for(n, 256, kDwGrid);
    uint32_t rpb = (n + nb - 1) / nb;
    rpb = (rpb + 31) / 32 * 32;
    const uint32_t width = K * N + (BIAS ? N : 0);
    hipLaunchKernelGGL((gemm_dw_kernel<K, K0, N, BIAS>), dim3(nb), dim3(256), 0, s, n, rpb, x0, ld0, x1, ld1, dy,
                       lddy, partial);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((width + 255) / 256), dim3(256), 0, s, partial, nb, width,
                       width, dst);
}

}  // namespace

// ===========================================================================
// Host side
// ===========================================================================

void mpnn_prepare(MpnnGraph& g, hipStream_t s, bool transpose) {
    const uint32_t n = g.n;
    g.sd.resize(n ? n : 1);
    g.dpos.resize(n ? n : 1);
    g.deg.resize(n ? n : 1);
    if (g.dim != 2 && g.dim != 3) throw std::invalid_argument("graph dimension must be 2 or 3");
    g.nfeat.resize((size_t)kNodeFeats * (n ? n : 1));
    g.efeat.resize(g.nnz ? g.nnz : 1);
    g.tr.resize(g.nnz ? g.nnz : 1);
    g.scal.resize(8);
    if (!n) return;
    const unsigned nb = blocks_for(n, 256, 1024);
    g.part.resize((size_t)kPrepVals * nb);
    hipLaunchKernelGGL(prep_rows_kernel, dim3(nb), dim3(256), 0, s, n, g.row, g.col, g.val, g.xy, g.xy_stride, g.dim,
                       g.sd.p, g.dpos.p, g.deg.p, g.part.p);
    hipLaunchKernelGGL(prep_fold_kernel, dim3(1), dim3(kFoldThreads), 0, s, nb, g.part.p, g.scal.p);
    const uint32_t lanes = (uint64_t)g.nnz > 10ull * n ? 16u : 8u;  // ~15 entries per row in 3D, ~7 in 2D
    hipLaunchKernelGGL(prep_feat_kernel, dim3(blocks_for(n, 256 / lanes, 65536)), dim3(256), 0, s, n, lanes, g.row,
                       g.col, g.val,
                       g.xy, g.xy_stride, g.dim, g.bnd, g.sd.p, g.dpos.p, g.scal.p, g.nfeat.p, g.efeat.p,
                       transpose ? g.tr.p : nullptr, ctx().d_flags);
    g.has_transpose = transpose;
}

void MpnnNet::set_params(const float* host, hipStream_t s) {
    const uint32_t np = kParams.total;
    w.resize(np);
    m.resize(np);
    v.resize(np);
    grad.resize(np);
    wb.resize(np);
    wt.resize(kWtTotal);
    h2d(w.p, host, np);
    fill<float>(m.p, np, 0.f, s);
    fill<float>(v.p, np, 0.f, s);
    fill<float>(grad.p, np, 0.f, s);
    adam_t = 0;
    hipLaunchKernelGGL(to_bf16_kernel, dim3((np + 255) / 256), dim3(256), 0, s, np, w.p, wb.p);
    hipLaunchKernelGGL(transpose_weights_kernel, dim3((kWtTotal + 255) / 256), dim3(256), 0, s, wb.p, wt.p);
    weights_set = true;
}

void MpnnNet::get_params(float* host) {
    if (!weights_set) throw std::invalid_argument("network weights not set");
    d2h(host, w.p, kParams.total);
}

void MpnnNet::reserve(uint32_t n, uint32_t nnz) {
    n = n ? n : 1;
    nnz = nnz ? nnz : 1;
    if (n > cap_n) {
        for (auto& b : H) b.resize((size_t)n * kHid);
        for (auto& b : PQ) b.resize((size_t)n * kPq);
        for (auto& b : M) b.resize((size_t)n * kHid);
        for (auto& b : U) b.resize((size_t)n * kHid);
        RSD.resize((size_t)n * kRsd);
        od.resize(n);
        dRSD.resize((size_t)n * kRsd);
        dHa.resize((size_t)n * kHid);
        dHb.resize((size_t)n * kHid);
        dX.resize((size_t)n * kPq);
        dPQ.resize((size_t)n * kPq);
        dU.resize((size_t)n * kHid);
        cap_n = n;
    }
    if (nnz > cap_nnz) {
        gval.resize(nnz);
        y.resize(nnz);
        cap_nnz = nnz;
    }
    const size_t pw = (size_t)kDwGrid * (3 * kHid * kHid + kHid);
    const size_t pe = (size_t)kEdgeGrid * kHeadPart;
    partials.resize(pw > pe ? pw : pe);
    lpart.resize(2ull * 4096);
    lsc.resize(4);
}

void MpnnNet::forward(const MpnnGraph& g, hipStream_t s, bool keep_pre) {
    if (!weights_set) throw std::invalid_argument("network weights not set");
    const uint32_t n = g.n;
    reserve(n, g.nnz);
    if (!n) return;
    const unsigned nodes_grid = blocks_for(n, 4, 65536);
    hipLaunchKernelGGL(enc_fwd_kernel, dim3(nodes_grid), dim3(256), 0, s, n, g.nfeat.p, wb.p, w.p, H[0].p);
    for (uint32_t l = 0; l < kLayers; ++l) {
        run_gemm<64, 64, 128, kEpiStore>(gargs(n, H[l].p, kHid, wt.p + wt_pq(l), PQ[l].p, kPq), s);
        hipLaunchKernelGGL(msg_fwd_kernel, dim3(nodes_grid), dim3(256), 0, s, n, g.row, g.col, g.efeat.p,
                           g.deg.p, PQ[l].p, wb.p, w.p, l, M[l].p);
        GemmArgs a = gargs(n, H[l].p, kHid, wt.p + wt_u(l), H[l + 1].p, kHid);
        a.x1 = M[l].p;
        a.ld1 = kHid;
        a.bias = w.p + kParams.u_b[l];
        a.add0 = H[l].p;
        a.lda0 = kHid;
        a.u = keep_pre ? U[l].p : nullptr;
        run_gemm<128, 64, 64, kEpiResRelu>(a, s);
    }
    run_gemm<64, 64, 192, kEpiStore>(gargs(n, H[kLayers].p, kHid, wt.p + kWtDec, RSD.p, kRsd), s);
    hipLaunchKernelGGL(head_fwd_kernel, dim3(nodes_grid), dim3(256), 0, s, n, g.row, g.col, g.val, g.dpos.p, g.bnd,
                       g.efeat.p, RSD.p, wb.p, w.p, gval.p, od.p);
}

void MpnnNet::loss(const MpnnGraph& g, hipStream_t s) {
    const uint32_t n = g.n;
    if (!n) return;
    const unsigned nb = blocks_for(n, 256, 4096);
    hipLaunchKernelGGL(loss_kernel, dim3(nb), dim3(256), 0, s, n, g.row, g.col, g.val, g.sd.p, g.dpos.p, gval.p, y.p,
                       lpart.p, ctx().d_flags);
    hipLaunchKernelGGL(loss_fold_kernel, dim3(1), dim3(64), 0, s, nb, n, lpart.p, lsc.p);
}

void MpnnNet::backward(const MpnnGraph& g, hipStream_t s) {
    const uint32_t n = g.n;
    if (!n) return;
    const unsigned eg = blocks_for(n, 4, kEdgeGrid);
    float* G = grad.p;
    hipLaunchKernelGGL(head_bwd_kernel, dim3(eg), dim3(256), 0, s, n, g.row, g.col, g.val, g.dpos.p, g.tr.p, g.bnd,
                       g.efeat.p, RSD.p, wb.p, w.p, gval.p, od.p, y.p, lsc.p, dRSD.p, partials.p);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((kHeadPart + 255) / 256), dim3(256), 0, s, partials.p, eg,
                       kHeadPart, kHeadPart, G + kParams.de_w);
    run_dw<64, 64, 192, false>(n, H[kLayers].p, kHid, nullptr, 0, dRSD.p, kRsd, partials.p, G + kParams.dec_w, s);
    // dH3 = dRSD W_dec^T
    run_gemm<192, 192, 64, kEpiAdd>(gargs(n, dRSD.p, kRsd, wb.p + kParams.dec_w, dHa.p, kHid), s);
    float* dh = dHa.p;
    float* dh_next = dHb.p;
    for (int l = (int)kLayers - 1; l >= 0; --l) {
        const size_t tot = (size_t)n * kHid;
        hipLaunchKernelGGL(relu_mask_kernel, dim3(blocks_for(tot, 256, 65536)), dim3(256), 0, s, tot, dh, U[l].p,
                           dU.p);
        run_dw<128, 64, 64, true>(n, H[l].p, kHid, M[l].p, kHid, dU.p, kHid, partials.p, G + kParams.u_w[l], s);
        // dX = dU W_u^T  ([n][128]: dH part, dM part)
        run_gemm<64, 64, 128, kEpiStore>(gargs(n, dU.p, kHid, wb.p + kParams.u_w[l], dX.p, kPq), s);
        hipLaunchKernelGGL(msg_bwd_kernel, dim3(eg), dim3(256), 0, s, n, g.row, g.col, g.efeat.p, g.deg.p,
                           PQ[l].p, dX.p, wb.p, w.p, (uint32_t)l, dPQ.p, partials.p);
        hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, s, partials.p, eg, kMsgPart, kMsgPart,
                           G + kParams.e_w[l]);
        run_dw<64, 64, 128, false>(n, H[l].p, kHid, nullptr, 0, dPQ.p, kPq, partials.p, G + kParams.pq_w[l], s);
        // dH_l = dH_{l+1} + dX[:, :64] + dPQ W_pq^T
        GemmArgs a = gargs(n, dPQ.p, kPq, wb.p + kParams.pq_w[l], dh_next, kHid);
        a.add0 = dh;
        a.lda0 = kHid;
        a.add1 = dX.p;
        a.lda1 = kPq;
        run_gemm<128, 128, 64, kEpiAdd>(a, s);
        float* t = dh;
        dh = dh_next;
        dh_next = t;
    }
    const unsigned encg = blocks_for(n, 4, kEdgeGrid);
    hipLaunchKernelGGL(enc_bwd_kernel, dim3(encg), dim3(256), 0, s, n, g.nfeat.p, H[0].p, dh, partials.p);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((kEncPart + 255) / 256), dim3(256), 0, s, partials.p, encg,
                       kEncPart, kEncPart, G + kParams.enc_w);
}

void MpnnNet::adam(float lr, float b1, float b2, float eps, hipStream_t s) {
    const uint32_t np = kParams.total;
    ++adam_t;
    const float c1 = 1.f - powf(b1, (float)adam_t), c2 = 1.f - powf(b2, (float)adam_t);
    hipLaunchKernelGGL(adam_kernel, dim3((np + 255) / 256), dim3(256), 0, s, np, w.p, m.p, v.p, grad.p, lr, b1, b2,
                       eps, c1, c2, wb.p);
    hipLaunchKernelGGL(transpose_weights_kernel, dim3((kWtTotal + 255) / 256), dim3(256), 0, s, wb.p, wt.p);
}

void MpnnNet::pack(const MpnnGraph& g, LowerFactor<float>& L, hipStream_t s) {
    if (!g.n) return;
    hipLaunchKernelGGL(pack_kernel, dim3(grid_for(g.n, 256)), dim3(256), 0, s, g.n, g.row, g.col, g.sd.p, g.dpos.p,
                       gval.p, L.slot.p, L.up.val.p, L.lo.val.p, L.diag.p, ctx().d_flags);
}

void mpnn_check_flags() {
    const uint32_t fl = read_scalar(ctx().d_flags + kFlagError);
    if (fl) {
        fill<uint32_t>(ctx().d_flags, 1, 0u, ctx().stream);
        throw std::runtime_error("device error flag raised in the learned preconditioner");
    }
}

}  // namespace stream_amd

using namespace stream_amd;

// ===========================================================================
// C-ABI (stream_learned.h).  No reference counterpart: the reference's only
// preconditioner is PrecJacobi (include/numerics/preconditioner.hpp:60-72);
// the handle follows its create/destroy pattern.
// ===========================================================================

extern "C" {

PrecLearned* PrecLearned_create(void) {
    try {
        (void)ctx();
        return new PrecLearned();
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return nullptr;
    }
}

void PrecLearned_destroy(PrecLearned* h) { delete h; }

uint32_t PrecLearned_param_count(void) { return kParams.total; }

int32_t PrecLearned_set_params(PrecLearned* h, const float* w, uint32_t count) {
    return guarded([&] {
        if (!h || !w || count != kParams.total) throw std::invalid_argument("PrecLearned_set_params: bad arguments");
        h->net.set_params(w, ctx().stream);
        SA_HIP(hipStreamSynchronize(ctx().stream));
    });
}

int32_t PrecLearned_get_params(PrecLearned* h, float* w, uint32_t count) {
    return guarded([&] {
        if (!h || !w || count != kParams.total) throw std::invalid_argument("PrecLearned_get_params: bad arguments");
        h->net.get_params(w);
    });
}

int32_t PrecLearned_set_adam(PrecLearned* h, float b1, float b2, float eps) {
    return guarded([&] {
        if (!h || !(b1 >= 0.f && b1 < 1.f) || !(b2 >= 0.f && b2 < 1.f) || !(eps > 0.f))
            throw std::invalid_argument("PrecLearned_set_adam: bad arguments");
        h->b1 = b1;
        h->b2 = b2;
        h->eps = eps;
    });
}

// Host system: CSR of A (FP64, sorted columns, symmetric pattern), point
// coordinates [n][dim] in the matrix's row order and the Dirichlet flags.
int32_t PrecLearned_set_system_dim(PrecLearned* h, uint32_t dim, uint32_t n, const uint32_t* row,
                                   const uint32_t* col, const double* val, const double* xyz, const uint8_t* bnd) {
    return guarded([&] {
        if (!h || !row || (n && (!xyz || !bnd))) throw std::invalid_argument("PrecLearned_set_system: null argument");
        if (dim != 2 && dim != 3) throw std::invalid_argument("PrecLearned_set_system: dim must be 2 or 3");
        const uint32_t nnz = row[n];
        if (nnz && (!col || !val)) throw std::invalid_argument("PrecLearned_set_system: null argument");
        // the kernels rely on ascending columns within a row (the upper part
        // follows the diagonal; transpose positions by binary search)
        for (uint32_t i = 0; i < n; ++i) {
            if (row[i] > row[i + 1]) throw std::invalid_argument("PrecLearned_set_system: row pointers decrease");
            for (uint32_t e = row[i] + 1; e < row[i + 1]; ++e)
                if (col[e] <= col[e - 1])
                    throw std::invalid_argument("PrecLearned_set_system: columns of a row must be ascending");
        }
        MpnnGraph& g = h->graph;
        g.own_row.resize(n + 1);
        g.own_col.resize(nnz ? nnz : 1);
        g.own_val.resize(nnz ? nnz : 1);
        g.own_xy.resize((size_t)dim * (n ? n : 1));
        g.own_bnd.resize(n ? n : 1);
        h2d(g.own_row.p, row, n + 1);
        h2d(g.own_col.p, col, nnz);
        h2d(g.own_val.p, val, nnz);
        h2d(g.own_xy.p, xyz, (size_t)dim * n);
        h2d(g.own_bnd.p, bnd, n);
        g.n = n;
        g.nnz = nnz;
        g.row = g.own_row.p;
        g.col = g.own_col.p;
        g.val = g.own_val.p;
        g.xy = g.own_xy.p;
        g.xy_stride = dim;
        g.dim = dim;
        g.bnd = g.own_bnd.p;
        mpnn_prepare(g, ctx().stream);
        mpnn_check_flags();
    });
}
int32_t PrecLearned_set_system(PrecLearned* h, uint32_t n, const uint32_t* row, const uint32_t* col,
                               const double* val, const double* xy, const uint8_t* bnd) {
    return PrecLearned_set_system_dim(h, 2, n, row, col, val, xy, bnd);
}

// Forward pass and loss on the bound system.  loss: Kaporin loss (nullable);
// ghat: Ghat per CSR entry [nnz] (upper entries and diagonal; lower entries
// are not written), nullable.
int32_t PrecLearned_forward(PrecLearned* h, double* loss, float* ghat) {
    return guarded([&] {
        if (!h) throw std::invalid_argument("PrecLearned_forward: null handle");
        hipStream_t s = ctx().stream;
        h->net.forward(h->graph, s);
        h->net.loss(h->graph, s);
        mpnn_check_flags();
        if (loss) *loss = read_scalar(h->net.lsc.p + 2);
        if (ghat && h->graph.nnz) d2h(ghat, h->net.gval.p, h->graph.nnz);
    });
}

// Loss and its gradient with respect to the FP32 master parameters [P].
int32_t PrecLearned_gradient(PrecLearned* h, double* loss, float* grad) {
    return guarded([&] {
        if (!h) throw std::invalid_argument("PrecLearned_gradient: null handle");
        if (!h->graph.has_transpose)
            throw std::invalid_argument("the bound system was prepared for inference only: bind it for training (PrecLearned_bind_step / PrecLearned_set_system)");
        hipStream_t s = ctx().stream;
        h->net.forward(h->graph, s);
        h->net.loss(h->graph, s);
        h->net.backward(h->graph, s);
        mpnn_check_flags();
        if (loss) *loss = read_scalar(h->net.lsc.p + 2);
        if (grad) d2h(grad, h->net.grad.p, kParams.total);
    });
}

// Average duration (ms, HIP events on the library stream over `reps`
// launches) of the layer-0 [P|Q] = H W_pq GEMM on the bound system, and its
// FLOP count per launch (2 n 64 128).
int32_t PrecLearned_time_gemm(PrecLearned* h, uint32_t reps, double* ms, double* flops) {
    return guarded([&] {
        if (!h || !ms || !reps) throw std::invalid_argument("PrecLearned_time_gemm: bad arguments");
        hipStream_t s = ctx().stream;
        MpnnNet& net = h->net;
        const uint32_t n = h->graph.n;
        net.forward(h->graph, s);
        hipEvent_t e0, e1;
        SA_HIP(hipEventCreate(&e0));
        SA_HIP(hipEventCreate(&e1));
        const GemmArgs a = gargs(n, net.H[0].p, kHid, net.wt.p + wt_pq(0), net.PQ[0].p, kPq);
        run_gemm<64, 64, 128, kEpiStore>(a, s);
        SA_HIP(hipEventRecord(e0, s));
        for (uint32_t k = 0; k < reps; ++k) run_gemm<64, 64, 128, kEpiStore>(a, s);
        SA_HIP(hipEventRecord(e1, s));
        SA_HIP(hipEventSynchronize(e1));
        float t = 0.f;
        SA_HIP(hipEventElapsedTime(&t, e0, e1));
        SA_HIP(hipEventDestroy(e0));
        SA_HIP(hipEventDestroy(e1));
        *ms = (double)t / reps;
        if (flops) *flops = 2.0 * n * kHid * kPq;
    });
}

// One Adam step on the bound system.  loss (nullable) is the loss before
// the update; reading it synchronises.
int32_t PrecLearned_train_step(PrecLearned* h, float lr, double* loss) {
    return guarded([&] {
        if (!h || !(lr > 0.f)) throw std::invalid_argument("PrecLearned_train_step: bad arguments");
        if (!h->graph.has_transpose)
            throw std::invalid_argument("the bound system was prepared for inference only: bind it for training (PrecLearned_bind_step / PrecLearned_set_system)");
        hipStream_t s = ctx().stream;
        h->net.forward(h->graph, s);
        h->net.loss(h->graph, s);
        h->net.backward(h->graph, s);
        h->net.adam(lr, h->b1, h->b2, h->eps, s);
        if (loss) {
            mpnn_check_flags();
            *loss = read_scalar(h->net.lsc.p + 2);
        }
    });
}

}  // extern "C"
