// Device value types and small geometric helpers shared by the LBVH and the
// alpha-shape kernels.
//
// Sph4<S>: a sphere (centre, radius) padded to four scalars so that one
// 16 B (float) or 2 x 16 B (double) vector load fetches it; 2D spheres leave
// c[2] unused.  This is the device form of the reference's Sphere<S,dim>
// (include/geometry/primitives.hpp:123-142).  BoxD<S,D> is BBox<S,dim>
// (include/geometry/bbox.hpp:48-121) in lo/hi form.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace stream_amd {

template <class S> struct alignas(4 * sizeof(S)) Sph4 {
    S c[3];
    S r;
};

template <class S, int D> struct BoxD {
    S lo[D];
    S hi[D];
};

template <class S> __host__ __device__ constexpr S max_finite();
template <> __host__ __device__ constexpr float max_finite<float>() { return FLT_MAX; }
template <> __host__ __device__ constexpr double max_finite<double>() { return DBL_MAX; }

template <class S, int D> __host__ __device__ inline BoxD<S, D> box_empty() {
    BoxD<S, D> b;
    for (int a = 0; a < D; ++a) {
        b.lo[a] = max_finite<S>();
        b.hi[a] = -max_finite<S>();
    }
    return b;
}
// combineBox: per-axis fmin / fmax (bbox.hpp:88-96)
template <class S, int D> __host__ __device__ inline void box_combine(BoxD<S, D>& a, const BoxD<S, D>& b) {
    for (int k = 0; k < D; ++k) a.lo[k] = fmin(a.lo[k], b.lo[k]);
    for (int k = 0; k < D; ++k) a.hi[k] = fmax(a.hi[k], b.hi[k]);
}
// Sphere::get_bbox = [c - r, c + r] (primitives.hpp:136-141)
template <class S, int D> __host__ __device__ inline BoxD<S, D> sphere_box(const Sph4<S>& s) {
    BoxD<S, D> b;
    for (int a = 0; a < D; ++a) {
        b.lo[a] = s.c[a] - s.r;
        b.hi[a] = s.c[a] + s.r;
    }
    return b;
}
// Box of the centre only (DirectAlphaShape2D.hpp:87-99)
template <class S, int D> __host__ __device__ inline BoxD<S, D> centre_box(const Sph4<S>& s) {
    BoxD<S, D> b;
    for (int a = 0; a < D; ++a) b.lo[a] = b.hi[a] = s.c[a];
    return b;
}

__device__ __forceinline__ int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

// Bit interleaving ("part-1-by-1" / "part-1-by-2" magic masks).
__host__ __device__ inline uint64_t spread2(uint32_t v) {
    uint64_t x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}
__host__ __device__ inline uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x001F00000000FFFFull;
    x = (x | (x << 16)) & 0x001F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// Coherent exchange of small POD records between workgroups running on
// different XCDs: every 64-bit word is written and read with a relaxed
// agent-scope atomic (write-through / read-through the non-coherent L2),
// and the ordering with the handshake atomic is enforced by waiting for
// outstanding vector-memory operations (the same scheme as rocPRIM's
// look-back state).
template <class T> __device__ __forceinline__ void coherent_store(T* p, const T& v) {
    static_assert(sizeof(T) % 8 == 0, "record must be a multiple of 8 bytes");
    const uint64_t* src = reinterpret_cast<const uint64_t*>(&v);
    uint64_t* dst = reinterpret_cast<uint64_t*>(p);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 8); ++k)
        __hip_atomic_store(dst + k, src[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ T coherent_load(const T* p) {
    static_assert(sizeof(T) % 8 == 0, "record must be a multiple of 8 bytes");
    T v;
    uint64_t* dst = reinterpret_cast<uint64_t*>(&v);
    const uint64_t* src = reinterpret_cast<const uint64_t*>(p);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 8); ++k)
        dst[k] = __hip_atomic_load(src + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return v;
}
// Wait until this thread's vector-memory operations have completed.
__device__ __forceinline__ void vmem_release_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));
}
__device__ __forceinline__ void vmem_acquire_order() {
    __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

}  // namespace stream_amd
