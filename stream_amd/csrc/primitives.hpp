// Wave64 building blocks and device-wide primitives for gfx950:
//   * wave/block scans built on DPP-backed shuffles and ballots,
//   * a single-pass exclusive scan with decoupled look-back,
//   * a onesweep LSD radix sort of (u64 key, u32 value) pairs: one histogram
//     pass over all digits, then one pass per 8-bit digit that ranks keys
//     with wave ballot matching, looks back for the digit prefix of earlier
//     tiles and scatters through LDS so that writes stay coalesced.
//
// These replace AVA's scan::inclusive_sum / inplace_inclusive_sum and
// sort::sort_pairs (call sites: include/geometry/lbvh.hpp:217-235,
// src/mesh/DirectAlphaShape2D.cpp:228-281,604-636).  The sort is stable, so
// equal keys keep their input order, as a stable sort_pairs does.
#pragma once
#include "runtime.hpp"

namespace stream_amd {

constexpr int kWave = 64;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, kWave);
        if (lane >= (uint32_t)o) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Exclusive scan of one value per thread over a block of NT threads.
// s_tmp must hold NT/64 words.  Returns the exclusive prefix; *total gets
// the block sum.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
    constexpr int NW = NT / kWave;
    const uint32_t lane = lane_id();
    const uint32_t w = threadIdx.x / kWave;
    const uint32_t incl = wave_incl_scan(v);
    if (lane == kWave - 1) s_tmp[w] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        const uint32_t t = s_tmp[k];
        off += ((uint32_t)k < w) ? t : 0u;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return off + incl - v;
}

// Relaxed agent-scope loads and stores of look-back status words.  They
// carry no ordering: each status word packs its flag and its value into
// one word, so a reader that sees the flag sees the value with it.  Data
// published in any other word needs real release/acquire ordering.
__device__ __forceinline__ uint32_t ld_relaxed_u32(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed_u32(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_relaxed_u64(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed_u64(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Spin bound of every look-back wait: a wait that runs out raises
// kErrSpin in the context flags instead of hanging the device.
constexpr uint32_t kSpinLimit = 1u << 24;

// Index of the error word in Context::d_flags.
constexpr int kFlagError = 0;

__device__ __forceinline__ void raise_flag(uint32_t* flags, uint32_t code) {
    atomicMax(&flags[kFlagError], code);
}

// ---------------------------------------------------------------------------
// Host launchers (primitives.hip).
// ---------------------------------------------------------------------------

// Exclusive scan: out[i] = sum(in[0..i)), out[n] = total.  in is u8 or u32.
// Workspace is owned by the caller (see scan_workspace_words).
size_t scan_workspace_words(uint32_t n);
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint64_t* ws, hipStream_t s);
void exclusive_scan_u8(const uint8_t* in, uint32_t* out, uint32_t n, uint64_t* ws, hipStream_t s);

// Radix sort of (u64 key, u32 value) pairs over the low `bits` key bits
// (multiple of 8, <= 64).  Keys/values are sorted from (k0,v0) and the
// result lands in (k0,v0) or (k1,v1): the function returns true when the
// sorted data is in (k1,v1).  hist must already hold the per-digit
// histograms (8 x 256 u32, digit d of pass p at hist[p*256+d]) unless
// compute_hist is set.
struct RadixWork {
    uint32_t* hist = nullptr;      // [8][256]
    uint32_t* status = nullptr;    // [passes][tiles][256]
    uint32_t* counters = nullptr;  // [8]
    uint32_t tiles = 0;
};
constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
size_t radix_workspace_words(uint32_t n);
RadixWork radix_work_from(uint32_t* ws, uint32_t n);
// Zero histograms, tile counters and look-back status (fill kernel).
void radix_reset(const RadixWork& w, int passes, hipStream_t s);
void radix_histogram(const uint64_t* keys, uint32_t n, const RadixWork& w, hipStream_t s);
bool radix_sort_pairs(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, uint32_t n, int bits,
                      const RadixWork& w, uint32_t* flags, hipStream_t s);

// Grid-stride fill kernel (used instead of memset so the sequence may be
// captured in a graph).
template <class T> __global__ void fill_kernel(T* p, size_t n, T v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
template <class T> void fill(T* p, size_t n, T v, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(fill_kernel<T>, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, p, n, v);
}

// Field permutation between input order and Morton order (the reference's
// permuteForward / permuteBackward, include/mesh/DirectAlphaShape2D.hpp:
// 220-242): forward out[i] = in[perm[i]], backward out[perm[i]] = in[i],
// perm[i] = input index of Morton slot i.  Elements are elem_bytes wide;
// in and out are device buffers of n elements that do not overlap.
void permute_elements(const void* in, void* out, const uint32_t* perm, uint32_t n, uint32_t elem_bytes, bool forward,
                      hipStream_t s);
// The same on host arrays (copied to the device and back; synchronises).
void permute_host_elements(const void* in, void* out, const uint32_t* d_perm, uint32_t n, uint32_t elem_bytes,
                           bool forward);

// Host side: read the context's error word (synchronises the stream); when
// set, clear it and throw.
inline void check_error_flag() {
    const uint32_t f = read_scalar(ctx().d_flags + kFlagError);
    if (f != 0) {
        fill<uint32_t>(ctx().d_flags, 1, 0u, ctx().stream);
        throw std::runtime_error("device error flag raised (code " + std::to_string(f) + ")");
    }
}

}  // namespace stream_amd
