// Device-wide scan and onesweep radix sort (see primitives.hpp).
#include <mutex>
#include <vector>

#include "primitives.hpp"

namespace stream_amd {

// ---------------------------------------------------------------------------
// Single-pass exclusive scan, decoupled look-back.
// Tile = 256 threads x 8 items.  Status word per tile: flag in the high 32
// bits (1 = tile aggregate, 2 = inclusive prefix), value in the low 32.
// ws[0] is the tile ticket counter, ws[1 + t] the status of tile t.
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;
constexpr uint64_t kAgg = 1ull << 32;
constexpr uint64_t kInc = 2ull << 32;

size_t scan_workspace_words(uint32_t n) { return 1 + (size_t)(n + kScanTile - 1) / kScanTile; }

template <class Tin>
__global__ __launch_bounds__(kScanThreads) void scan_kernel(const Tin* __restrict__ in, uint32_t* __restrict__ out,
                                                            uint32_t n, uint64_t* ws, uint32_t* flags) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_tmp[kScanThreads / kWave];
    __shared__ uint32_t s_prefix;
    uint64_t* status = ws + 1;
    if (threadIdx.x == 0) s_tile = (uint32_t)atomicAdd((unsigned long long*)ws, 1ull);
    __syncthreads();
    const uint32_t tile = s_tile;
    const size_t base = (size_t)tile * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const size_t idx = base + k;
        v[k] = idx < n ? (uint32_t)in[idx] : 0u;
        sum += v[k];
    }
    uint32_t total;
    const uint32_t excl = block_excl_scan<kScanThreads>(sum, s_tmp, &total);
    if (threadIdx.x < kWave) {
        const uint32_t lane = lane_id();
        uint32_t prefix = 0;
        if (tile == 0) {
            if (lane == 0) st_relaxed_u64(status, kInc | total);
        } else {
            if (lane == 0) st_relaxed_u64(status + tile, kAgg | total);
            int64_t j0 = (int64_t)tile - 1;
            uint32_t spins = 0;
            while (true) {
                const int64_t j = j0 - (int64_t)lane;
                const uint64_t sv = (j >= 0) ? ld_relaxed_u64(status + j) : kInc;
                const uint32_t fl = (uint32_t)(sv >> 32);
                if (__ballot(fl == 0)) {
                    if (++spins > kSpinLimit) {
                        if (lane == 0) raise_flag(flags, kErrSpin);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                const uint64_t inc = __ballot(fl == 2);
                uint32_t val = (uint32_t)sv;
                if (inc) {
                    const int k = __ffsll((unsigned long long)inc) - 1;
                    val = ((int)lane <= k) ? val : 0u;
                    prefix += wave_sum(val);
                    break;
                }
                prefix += wave_sum(val);
                j0 -= kWave;
            }
            if (lane == 0) st_relaxed_u64(status + tile, kInc | (uint64_t)(prefix + total));
        }
        if (lane == 0) s_prefix = prefix;
    }
    __syncthreads();
    uint32_t run = s_prefix + excl;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const size_t idx = base + k;
        if (idx < n) {
            out[idx] = run;
            run += v[k];
            if (idx == (size_t)n - 1) out[n] = run;
        }
    }
}

// ws: scan_workspace_words(n) words, cleared here by a fill kernel (no
// memset, so the sequence may be captured); flags: the error word a look-back
// that runs out of spins raises.
template <class Tin>
static void exclusive_scan(const Tin* in, uint32_t* out, uint32_t n, uint64_t* ws, uint32_t* flags, hipStream_t s) {
    if (n == 0) {
        fill<uint32_t>(out, 1, 0u, s);
        return;
    }
    const size_t words = scan_workspace_words(n);
    fill<uint64_t>(ws, words, 0ull, s);
    const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(scan_kernel<Tin>, dim3(tiles), dim3(kScanThreads), 0, s, in, out, n, ws, flags);
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint64_t* ws, hipStream_t s) {
    exclusive_scan<uint32_t>(in, out, n, ws, ctx().d_flags, s);
}
void exclusive_scan_u8(const uint8_t* in, uint32_t* out, uint32_t n, uint64_t* ws, hipStream_t s) {
    exclusive_scan<uint8_t>(in, out, n, ws, ctx().d_flags, s);
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free scan entry point: the look-back workspace
// of up to kNoAllocScanMax items (ticket + one status word per 2048-item
// tile) and, in the last word, the error flag its kernels raise.  One copy
// per device, held until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns for the current device.
// ---------------------------------------------------------------------------
constexpr uint32_t kNoAllocScanMax = 1u << 28;
constexpr size_t kNoAllocScanWords = 1 + (size_t)(kNoAllocScanMax + kScanTile - 1) / kScanTile + 1;
static __device__ __align__(256) uint64_t scratch[kNoAllocScanWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint64_t* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the
// first call for that device.
uint64_t* scan_scratch() {
    int dev = 0;
    SA_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* ptr = nullptr;
        SA_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw std::runtime_error("hipGetSymbolAddress returned no address for the scan scratch");
        g_scratch_dev[dev] = (uint64_t*)ptr;
    }
    return g_scratch_dev[dev];
}
}  // namespace

// ---------------------------------------------------------------------------
// Onesweep radix sort.
// ---------------------------------------------------------------------------
size_t radix_workspace_words(uint32_t n) {
    const size_t tiles = (n + kSortTile - 1) / kSortTile;
    return 8 * 256 + 8 + 8 * tiles * 256;
}

RadixWork radix_work_from(uint32_t* ws, uint32_t n) {
    RadixWork w;
    w.tiles = (n + kSortTile - 1) / kSortTile;
    w.hist = ws;
    w.counters = ws + 8 * 256;
    w.status = ws + 8 * 256 + 8;
    return w;
}

void radix_reset(const RadixWork& w, int passes, hipStream_t s) {
    // hist, counters and the status of the passes that will run are contiguous
    fill<uint32_t>(w.hist, 8 * 256 + 8 + (size_t)passes * w.tiles * 256, 0u, s);
}

__global__ __launch_bounds__(256) void radix_hist_kernel(const uint64_t* __restrict__ keys, uint32_t n,
                                                         uint32_t* __restrict__ hist) {
    __shared__ uint32_t s_h[8 * 256];
    for (int i = threadIdx.x; i < 8 * 256; i += blockDim.x) s_h[i] = 0;
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
#pragma unroll
        for (int p = 0; p < 8; ++p) atomicAdd(&s_h[p * 256 + (uint32_t)((k >> (8 * p)) & 255u)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 8 * 256; i += blockDim.x)
        if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

void radix_histogram(const uint64_t* keys, uint32_t n, const RadixWork& w, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(radix_hist_kernel, dim3(grid_for(n, 256, 1024)), dim3(256), 0, s, keys, n, w.hist);
}

__global__ __launch_bounds__(kSortThreads) void radix_pass_kernel(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint64_t* __restrict__ kout,
    uint32_t* __restrict__ vout, uint32_t n, uint32_t shift, const uint32_t* __restrict__ hist,
    uint32_t* status, uint32_t* counter, uint32_t* flags) {
    constexpr int NW = kSortThreads / kWave;
    __shared__ uint64_t s_keys[kSortTile];
    __shared__ uint32_t s_vals[kSortTile];
    __shared__ uint32_t s_whist[NW][256];
    __shared__ uint32_t s_excl[256];
    __shared__ uint32_t s_dbase[256];
    __shared__ uint32_t s_tmp[NW];
    __shared__ uint32_t s_tile;

    const uint32_t t = threadIdx.x;
    const uint32_t lane = lane_id();
    const uint32_t w = t / kWave;
    if (t == 0) s_tile = atomicAdd(counter, 1u);
    for (int i = t; i < NW * 256; i += kSortThreads) (&s_whist[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t tile_start = tile * (uint32_t)kSortTile;
    const uint64_t lt = lanemask_lt();

    uint64_t key[kSortItems];
    uint32_t val[kSortItems];
    uint32_t rank[kSortItems];
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t idx = tile_start + w * (kSortItems * kWave) + k * kWave + lane;
        const bool valid = idx < n;
        key[k] = valid ? kin[idx] : 0ull;
        val[k] = valid ? vin[idx] : 0u;
        const uint32_t d = (uint32_t)(key[k] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint32_t bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const int leader = valid ? (__ffsll((unsigned long long)peers) - 1) : (int)lane;
        uint32_t old = 0;
        if (valid && (int)lane == leader) old = atomicAdd(&s_whist[w][d], (uint32_t)__popcll(peers));
        old = __shfl(old, leader, kWave);
        rank[k] = old + (uint32_t)__popcll(peers & lt);
    }
    __syncthreads();
    // per digit (thread t owns digit t): wave-exclusive offsets and tile count
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        const uint32_t x = s_whist[k][t];
        s_whist[k][t] = cnt;
        cnt += x;
    }
    uint32_t* my_status = status + (size_t)tile * 256 + t;
    st_relaxed_u32(my_status, ((tile == 0 ? 2u : 1u) << 30) | cnt);
    uint32_t tot;
    const uint32_t ex = block_excl_scan<kSortThreads>(cnt, s_tmp, &tot);
    const uint32_t gex = block_excl_scan<kSortThreads>(hist[t], s_tmp, &tot);
    s_excl[t] = ex;
    uint32_t prefix = 0;
    if (tile > 0) {
        int32_t j = (int32_t)tile - 1;
        uint32_t spins = 0;
        while (j >= 0) {
            const uint32_t sv = ld_relaxed_u32(status + (size_t)j * 256 + t);
            const uint32_t fl = sv >> 30;
            if (fl == 0) {
                if (++spins > kSpinLimit) {
                    raise_flag(flags, kErrSpin);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            prefix += sv & 0x3fffffffu;
            if (fl == 2) break;
            --j;
        }
        st_relaxed_u32(my_status, (2u << 30) | (prefix + cnt));
    }
    s_dbase[t] = gex + prefix - ex;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
        const uint32_t idx = tile_start + w * (kSortItems * kWave) + k * kWave + lane;
        if (idx < n) {
            const uint32_t d = (uint32_t)(key[k] >> shift) & 255u;
            const uint32_t pos = s_excl[d] + s_whist[w][d] + rank[k];
            s_keys[pos] = key[k];
            s_vals[pos] = val[k];
        }
    }
    __syncthreads();
    const uint32_t nvalid = (n - tile_start) < (uint32_t)kSortTile ? (n - tile_start) : (uint32_t)kSortTile;
    for (uint32_t s = t; s < nvalid; s += kSortThreads) {
        const uint64_t k = s_keys[s];
        const uint32_t d = (uint32_t)(k >> shift) & 255u;
        const uint32_t dest = s_dbase[d] + s;
        if (dest < n) {
            kout[dest] = k;
            vout[dest] = s_vals[s];
        } else {
            raise_flag(flags, kErrSpin);
        }
    }
}

bool radix_sort_pairs(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, uint32_t n, int bits,
                      const RadixWork& w, uint32_t* flags, hipStream_t s) {
    const int passes = (bits + 7) / 8;
    if (n <= 1 || passes == 0) return false;
    for (int p = 0; p < passes; ++p) {
        const bool even = (p % 2) == 0;
        hipLaunchKernelGGL(radix_pass_kernel, dim3(w.tiles), dim3(kSortThreads), 0, s, even ? k0 : k1,
                           even ? v0 : v1, even ? k1 : k0, even ? v1 : v0, n, (uint32_t)(8 * p), w.hist + 256 * p,
                           w.status + (size_t)p * w.tiles * 256, w.counters + p, flags);
    }
    return (passes % 2) == 1;
}

// one thread per word of an element; W words of type T per element
template <class T>
__global__ __launch_bounds__(256) void permute_words_kernel(const T* __restrict__ in, T* __restrict__ out,
                                                            const uint32_t* __restrict__ perm, uint32_t n, uint32_t W,
                                                            int forward) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = (uint64_t)n * W;
    if (t >= total) return;
    const uint64_t i = t / W, k = t - i * W;
    const uint64_t p = perm[i];
    if (forward)
        out[i * W + k] = in[p * W + k];
    else
        out[p * W + k] = in[i * W + k];
}

template <class T>
static void permute_words(const void* in, void* out, const uint32_t* perm, uint32_t n, uint32_t W, bool forward,
                          hipStream_t s) {
    const uint64_t total = (uint64_t)n * W;
    const uint64_t blocks = (total + 255) / 256;
    if (blocks > 0x7fffffffull) throw std::invalid_argument("permutation too large");
    hipLaunchKernelGGL((permute_words_kernel<T>), dim3((uint32_t)blocks), dim3(256), 0, s, (const T*)in, (T*)out,
                       perm, n, W, forward ? 1 : 0);
}

void permute_elements(const void* in, void* out, const uint32_t* perm, uint32_t n, uint32_t elem_bytes, bool forward,
                      hipStream_t s) {
    if (n == 0) return;
    if (elem_bytes == 0) throw std::invalid_argument("element size must be positive");
    if (in == out) throw std::invalid_argument("permutation in place is not supported (in == out)");
    // widest word that divides the element and both addresses
    const uintptr_t al = (uintptr_t)in | (uintptr_t)out;
    for (uint32_t w : {16u, 8u, 4u, 2u, 1u}) {
        if (elem_bytes % w || al % w) continue;
        const uint32_t W = elem_bytes / w;
        switch (w) {
            case 16: permute_words<uint4>(in, out, perm, n, W, forward, s); return;
            case 8: permute_words<uint2>(in, out, perm, n, W, forward, s); return;
            case 4: permute_words<uint32_t>(in, out, perm, n, W, forward, s); return;
            case 2: permute_words<uint16_t>(in, out, perm, n, W, forward, s); return;
            default: permute_words<uint8_t>(in, out, perm, n, W, forward, s); return;
        }
    }
}

void permute_host_elements(const void* in, void* out, const uint32_t* d_perm, uint32_t n, uint32_t elem_bytes,
                           bool forward) {
    if (n == 0) return;
    if (!in || !out) throw std::invalid_argument("null field array");
    if (elem_bytes == 0) throw std::invalid_argument("element size must be positive");
    const size_t bytes = (size_t)n * elem_bytes;
    DBuf<uint8_t> din, dout;
    din.resize(bytes);
    dout.resize(bytes);
    h2d(din.p, (const uint8_t*)in, bytes);
    permute_elements(din.p, dout.p, d_perm, n, elem_bytes, forward, ctx().stream);
    d2h((uint8_t*)out, dout.p, bytes);
}

__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

}  // namespace stream_amd

using namespace stream_amd;

extern "C" {

// Stable sort of (u64 key, u32 value) pairs over the low `bits` key bits,
// host arrays in and out.  Test and benchmark entry point for the sort that
// replaces ava::sort::sort_pairs (include/geometry/lbvh.hpp:217-235).
int32_t Stream_sort_pairs_u64(uint32_t n, const uint64_t* keys, const uint32_t* vals, uint64_t* keys_out,
                              uint32_t* vals_out, int32_t bits) {
    return guarded([&] {
        if (bits < 0 || bits > 64 || (bits % 8) != 0) throw std::invalid_argument("bits must be 0..64, multiple of 8");
        hipStream_t s = ctx().stream;
        DBuf<uint64_t> k0, k1;
        DBuf<uint32_t> v0, v1, ws;
        k0.resize(n);
        k1.resize(n);
        v0.resize(n);
        v1.resize(n);
        ws.resize(radix_workspace_words(n));
        h2d(k0.p, keys, n);
        h2d(v0.p, vals, n);
        RadixWork w = radix_work_from(ws.p, n);
        radix_reset(w, (bits + 7) / 8, s);
        radix_histogram(k0.p, n, w, s);
        const bool in1 = radix_sort_pairs(k0.p, v0.p, k1.p, v1.p, n, bits, w, ctx().d_flags, s);
        d2h(keys_out, in1 ? k1.p : k0.p, n);
        d2h(vals_out, in1 ? v1.p : v0.p, n);
        if (read_scalar(ctx().d_flags + kFlagError) != 0) throw std::runtime_error("device error flag raised");
    });
}

// Plain device copy of `nbytes` (multiple of 16) repeated `reps` times: the
// calibration kernel for hardware byte counters (its algorithmic traffic is
// exactly nbytes read + nbytes written per launch).  Returns the average ms.
int32_t Stream_copy_calibration(uint64_t nbytes, uint32_t reps, double* ms) {
    return guarded([&] {
        if (nbytes % 16 != 0 || nbytes == 0) throw std::invalid_argument("nbytes must be a positive multiple of 16");
        hipStream_t s = ctx().stream;
        DBuf<uint4> a, b;
        const size_t n = nbytes / 16;
        a.resize(n);
        b.resize(n);
        fill<uint4>(a.p, n, make_uint4(1, 2, 3, 4), s);
        hipEvent_t e0, e1;
        SA_HIP(hipEventCreate(&e0));
        SA_HIP(hipEventCreate(&e1));
        SA_HIP(hipEventRecord(e0, s));
        for (uint32_t k = 0; k < reps; ++k)
            hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, a.p, b.p, n);
        SA_HIP(hipEventRecord(e1, s));
        SA_HIP(hipEventSynchronize(e1));
        float t = 0;
        SA_HIP(hipEventElapsedTime(&t, e0, e1));
        SA_HIP(hipEventDestroy(e0));
        SA_HIP(hipEventDestroy(e1));
        *ms = reps ? (double)t / reps : 0.0;
    });
}

// Exclusive scan of u32 (host arrays): out[0..n] with out[n] = total.
int32_t Stream_exclusive_scan_u32(uint32_t n, const uint32_t* in, uint32_t* out) {
    return guarded([&] {
        hipStream_t s = ctx().stream;
        DBuf<uint32_t> din, dout;
        DBuf<uint64_t> ws;
        din.resize(n ? n : 1);
        dout.resize(n + 1);
        ws.resize(scan_workspace_words(n));
        h2d(din.p, in, n);
        exclusive_scan_u32(din.p, dout.p, n, ws.p, s);
        d2h(out, dout.p, n + 1);
        check_error_flag();
    });
}

// Exclusive scan of u8 counts into u32 offsets (host arrays).
int32_t Stream_exclusive_scan_u8(uint32_t n, const uint8_t* in, uint32_t* out) {
    return guarded([&] {
        hipStream_t s = ctx().stream;
        DBuf<uint8_t> din;
        DBuf<uint32_t> dout;
        DBuf<uint64_t> ws;
        din.resize(n ? n : 1);
        dout.resize(n + 1);
        ws.resize(scan_workspace_words(n));
        h2d(din.p, in, n);
        exclusive_scan_u8(din.p, dout.p, n, ws.p, s);
        d2h(out, dout.p, n + 1);
        check_error_flag();
    });
}

// Allocation-free exclusive scan of u32 on device buffers (no counterpart
// in the reference): out[0..n] with out[n] = total, in [n] and out [n + 1]
// device pointers the caller owns, n <= 2^28.  The look-back workspace is
// the file-scope `scratch` array (one per device; cleared by a fill kernel
// on every call, so a call never depends on what the last one left), so
// the call makes no device allocation.  Kernels run on `stream` (a
// hipStream_t; NULL = the null stream) and the call waits for them, then
// reads the error word.  Calls that overlap on one device share `scratch`
// and race.
int32_t Stream_exclusive_scan_u32_noalloc(uint32_t n, const uint32_t* d_in, uint32_t* d_out, void* stream) {
    return guarded([&] {
        if (n > kNoAllocScanMax)
            throw std::invalid_argument("Stream_exclusive_scan_u32_noalloc: n exceeds the scratch limit of 2^28 items");
        if (!d_out || (n > 0 && !d_in)) throw std::invalid_argument("Stream_exclusive_scan_u32_noalloc: null buffer");
        hipStream_t s = (hipStream_t)stream;
        uint64_t* ws = scan_scratch();
        uint32_t* flag = (uint32_t*)(ws + kNoAllocScanWords - 1);
        fill<uint32_t>(flag, 2, 0u, s);
        exclusive_scan<uint32_t>(d_in, d_out, n, ws, flag, s);
        SA_HIP(hipGetLastError());
        uint32_t f = 0;
        SA_HIP(hipMemcpyAsync(&f, flag, sizeof(f), hipMemcpyDeviceToHost, s));
        SA_HIP(hipStreamSynchronize(s));
        if (f != 0) throw std::runtime_error("Stream_exclusive_scan_u32_noalloc: device look-back ran out of spins");
    });
}

}  // extern "C"
