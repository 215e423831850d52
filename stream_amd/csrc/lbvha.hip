// Agglomerative LBVH (Apetrei 2014) on the GPU: the topology and the
// bottom-up refit in one pass.  Replaces include/geometry/lbvh_agglo.hpp
// (LBVHa::build, :62-361).
//
// Leaves are nodes 0..n-1 (Morton order), internal nodes n..2n-2; internal
// node k+n splits between leaves k and k+1.  Each leaf thread climbs: a node
// covering [imin, imax] attaches to its right parent (imax + n) when
// imin == 0 or (imax != n-1 and delta(imax) < delta(imin-1)), else to its
// left parent (imin - 1 + n), with delta(i) = m_i xor m_{i+1}; the second
// thread to reach a parent combines the two children's boxes and goes on.
// The topology depends only on the deltas, and box combination is min/max,
// so the result is deterministic.  Cross-XCD visibility of the child
// records uses the same relaxed-atomic + vmcnt ordering as the Karras refit.
#include <vector>

#include "lbvh.hpp"
#include "traversal.hpp"

namespace stream_amd {

namespace {

__device__ __forceinline__ void st_u32(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_u32(const uint32_t* p) {
    return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class S, int D>
__global__ __launch_bounds__(256) void agglo_init_kernel(const Sph4<S>* __restrict__ obj_m,
                                                         const uint64_t* __restrict__ codes, uint32_t n,
                                                         uint32_t leaf_mode, BoxD<S, D>* __restrict__ data,
                                                         uint64_t* __restrict__ delta, uint32_t* __restrict__ touched) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    data[i] = leaf_mode == kLeafCentre ? centre_box<S, D>(obj_m[i]) : sphere_box<S, D>(obj_m[i]);
    if (i + 1 < n) {
        delta[i] = codes[i] ^ codes[i + 1];
        touched[i] = 0;
    }
}

template <class S, int D>
__global__ __launch_bounds__(256) void agglo_build_kernel(uint32_t n, const uint64_t* __restrict__ delta,
                                                          uint32_t* children, uint32_t* range, uint32_t* parent_of,
                                                          uint32_t* split, uint32_t* touched, BoxD<S, D>* data,
                                                          uint32_t* root, uint32_t* flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t idx = i, imin = i, imax = i;
    BoxD<S, D> box = data[i];
    for (uint32_t level = 0;; ++level) {
        if (level > 2048u) {
            raise_flag(flags, kErrCapacity);
            break;
        }
        if (imin == 0 && imax == n - 1) {
            st_u32(&parent_of[idx], 2u * n);
            st_u32(root, idx);
            break;
        }
        const bool right_parent = imin == 0 || (imax != n - 1 && delta[imax] < delta[imin - 1]);
        const uint32_t parent = right_parent ? imax + n : imin + n - 1;
        const uint32_t k = parent - n;
        const uint32_t side = right_parent ? 0u : 1u;
        st_u32(&range[2 * k + side], right_parent ? imin : imax);
        st_u32(&children[2 * k + side], idx);
        st_u32(&parent_of[idx], parent);
        vmem_release_order();
        const uint32_t old = __hip_atomic_fetch_add(&touched[k], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!old) break;
        vmem_acquire_order();
        const uint32_t sibling = ld_u32(&children[2 * k + (1u - side)]);
        box_combine(box, coherent_load(&data[sibling]));
        coherent_store(&data[parent], box);
        st_u32(&split[k], right_parent ? imax : imin - 1);
        idx = parent;
        imin = ld_u32(&range[2 * k]);
        imax = ld_u32(&range[2 * k + 1]);
    }
}

// Sphere-sphere collision count over the agglomerative tree with leaf
// buckets of up to leaf_size objects (tests/geometry/
// lbvh_agglo_simple_collision.cpp:105-160).
template <class S, int D>
__global__ __launch_bounds__(256) void agglo_collide_kernel(const Sph4<S>* __restrict__ obj_m, uint32_t n,
                                                            const uint32_t* __restrict__ children,
                                                            const uint32_t* __restrict__ range,
                                                            const BoxD<S, D>* __restrict__ data,
                                                            const uint32_t* __restrict__ root, uint32_t leaf_size,
                                                            unsigned long long* __restrict__ total,
                                                            uint32_t* __restrict__ flags) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c = 0;
    if (tid < n) {
        uint32_t stack[64];
        uint32_t ss = 0;
        stack[ss++] = *root;
        const Sph4<S> q = obj_m[tid];
        for (uint32_t visits = 0; ss && visits < 2 * n; ++visits) {
            const uint32_t cur = stack[--ss];
            if (cur < n) {  // single-leaf tree
                continue;
            }
            for (uint32_t ch = 0; ch < 2; ++ch) {
                const uint32_t cid = children[2 * (cur - n) + ch];
                uint32_t rmin, rmax;
                if (cid >= n) {
                    rmin = range[2 * (cid - n)];
                    rmax = range[2 * (cid - n) + 1];
                } else {
                    rmin = rmax = cid;
                }
                const bool is_leaf = cid < n || (rmax - rmin < leaf_size);
                if (!is_leaf) {
                    const BoxD<S, D> b = data[cid];
                    bool in = true;
                    for (int a = 0; a < D; ++a) {
                        in &= q.c[a] - q.r < b.hi[a];
                        in &= q.c[a] + q.r > b.lo[a];
                    }
                    if (in) {
                        if (ss < 64)
                            stack[ss++] = cid;
                        else
                            raise_flag(flags, kErrCapacity);
                    }
                } else {
                    for (uint32_t o = rmin; o <= rmax; ++o) {
                        const Sph4<S> s = obj_m[o];
                        const S rs = q.r + s.r;
                        S d2 = 0;
                        for (int a = 0; a < D; ++a) {
                            const S dd = q.c[a] - s.c[a];
                            d2 += dd * dd;
                        }
                        if (d2 <= rs * rs && o != tid) ++c;
                    }
                }
            }
        }
    }
    c = wave_sum(c);
    if (lane_id() == 0 && c) atomicAdd(total, (unsigned long long)c);
}

template <class S, int D> struct Agglo {
    LbvhGpu<S, D> base;
    DBuf<BoxD<S, D>> data;  // [2n-1]
    DBuf<uint64_t> delta;   // [n-1]
    DBuf<uint32_t> children, range, parent_of, split, touched, root;

    void build(const Sph4<S>* d_obj, uint32_t n, int bits, uint32_t leaf_mode, hipStream_t s) {
        if (n < 2) throw std::invalid_argument("LBVHa needs at least 2 objects");
        base.sort_only(d_obj, n, bits, s);
        data.resize(2ull * n);
        delta.resize(n);
        children.resize(2ull * n);
        range.resize(2ull * n);
        parent_of.resize(2ull * n);
        split.resize(n);
        touched.resize(n);
        root.resize(1);
        hipLaunchKernelGGL((agglo_init_kernel<S, D>), dim3(grid_for(n, 256)), dim3(256), 0, s, base.obj_m.p,
                           base.codes_sorted, n, leaf_mode, data.p, delta.p, touched.p);
        hipLaunchKernelGGL((agglo_build_kernel<S, D>), dim3(grid_for(n, 256)), dim3(256), 0, s, n, delta.p,
                           children.p, range.p, parent_of.p, split.p, touched.p, data.p, root.p, ctx().d_flags);
    }
};

template <class S, int D> std::vector<Sph4<S>> pack_rows(uint32_t n, const S* xyr) {
    std::vector<Sph4<S>> h(n);
    for (uint32_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) h[i].c[a] = a < D ? xyr[(size_t)(D + 1) * i + a] : (S)0;
        h[i].r = xyr[(size_t)(D + 1) * i + D];
    }
    return h;
}

void check_device_flags() {
    const uint32_t f = read_scalar(ctx().d_flags + kFlagError);
    if (f) {
        fill<uint32_t>(ctx().d_flags, 1, 0u, ctx().stream);
        throw std::runtime_error("device error flag raised in the agglomerative LBVH");
    }
}

template <class S, int D>
int32_t agglo_build_host(uint32_t n, const S* xyr, int32_t bits, int32_t leaf_mode, uint32_t* perm,
                         uint32_t* children, uint32_t* range, uint32_t* parent, uint32_t* root, S* boxes,
                         uint32_t* split) {
    return guarded([&] {
        hipStream_t s = ctx().stream;
        auto h = pack_rows<S, D>(n, xyr);
        DBuf<Sph4<S>> d_obj;
        d_obj.resize(n ? n : 1);
        h2d(d_obj.p, h.data(), n);
        Agglo<S, D> t;
        t.build(d_obj.p, n, bits, (uint32_t)leaf_mode, s);
        check_device_flags();
        if (perm) d2h(perm, t.base.perm, n);
        if (children) d2h(children, t.children.p, 2ull * (n - 1));
        if (range) d2h(range, t.range.p, 2ull * (n - 1));
        if (parent) d2h(parent, t.parent_of.p, 2ull * n - 1);
        if (root) d2h(root, t.root.p, 1);
        if (boxes) d2h((BoxD<S, D>*)boxes, t.data.p, 2ull * n - 1);
        if (split) d2h(split, t.split.p, n - 1);
    });
}

template <class S, int D>
int32_t agglo_collide_host(uint32_t n, const S* xyr, int32_t bits, uint32_t leaf_size, uint64_t* total) {
    return guarded([&] {
        if (leaf_size < 1) throw std::invalid_argument("leaf_size must be >= 1");
        hipStream_t s = ctx().stream;
        auto h = pack_rows<S, D>(n, xyr);
        DBuf<Sph4<S>> d_obj;
        d_obj.resize(n ? n : 1);
        h2d(d_obj.p, h.data(), n);
        Agglo<S, D> t;
        t.build(d_obj.p, n, bits, kLeafSphere, s);
        DBuf<unsigned long long> tot;
        tot.resize(1);
        fill<unsigned long long>(tot.p, 1, 0ull, s);
        hipLaunchKernelGGL((agglo_collide_kernel<S, D>), dim3(grid_for(n, 256)), dim3(256), 0, s, t.base.obj_m.p, n,
                           t.children.p, t.range.p, t.data.p, t.root.p, leaf_size, tot.p, ctx().d_flags);
        *total = read_scalar(tot.p);
        check_device_flags();
    });
}

// Push n (idx, cost) items into a 32-slot TraversalMinHeap on the device and
// pop them all (one thread): the pop order, and how many pushes fit.
__global__ void heap_order_kernel(uint32_t n, const uint32_t* idx, const float* cost, uint32_t* out_idx,
                                  float* out_cost, int32_t* pushed) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    TraversalMinHeap<uint32_t, float, 32> h;
    int ok = 0;
    for (uint32_t k = 0; k < n; ++k) ok += h.push(idx[k], cost[k]) ? 1 : 0;
    *pushed = ok;
    int m = 0;
    while (!h.empty()) {
        const auto p = h.peek();
        h.pop();
        out_idx[m] = p.first;
        out_cost[m] = p.second;
        ++m;
    }
}

}  // namespace

}  // namespace stream_amd

using namespace stream_amd;

extern "C" {

// Agglomerative LBVH over n host spheres, rows (x, y[, z], r).  Outputs
// (host, any may be null): the sort permutation, children [n-1][2] and
// ranges [n-1][2] of the internal nodes (internal node k is node id k + n),
// parent [2n-1] (2n at the root), the root id, the node boxes [2n-1] as rows
// (min..., max...) (leaves first) and the split index [n-1].  leaf_mode as
// LBVH2D_build_*.  Test surface of include/geometry/lbvh_agglo.hpp:62-361,
// which exports no C symbol.
int32_t LBVHa2D_build_f32(uint32_t n, const float* xyr, int32_t bits, int32_t leaf_mode, uint32_t* perm,
                          uint32_t* children, uint32_t* range, uint32_t* parent, uint32_t* root, float* boxes,
                          uint32_t* split) {
    return agglo_build_host<float, 2>(n, xyr, bits, leaf_mode, perm, children, range, parent, root, boxes, split);
}
int32_t LBVHa2D_build_f64(uint32_t n, const double* xyr, int32_t bits, int32_t leaf_mode, uint32_t* perm,
                          uint32_t* children, uint32_t* range, uint32_t* parent, uint32_t* root, double* boxes,
                          uint32_t* split) {
    return agglo_build_host<double, 2>(n, xyr, bits, leaf_mode, perm, children, range, parent, root, boxes, split);
}
int32_t LBVHa3D_build_f64(uint32_t n, const double* xyzr, int32_t bits, int32_t leaf_mode, uint32_t* perm,
                          uint32_t* children, uint32_t* range, uint32_t* parent, uint32_t* root, double* boxes,
                          uint32_t* split) {
    return agglo_build_host<double, 3>(n, xyzr, bits, leaf_mode, perm, children, range, parent, root, boxes, split);
}
// Sphere-sphere collision count (self excluded) with leaf buckets of up to
// leaf_size objects (lbvh_agglo_simple_collision.cpp:100-165).
int32_t LBVHa2D_collisions_f32(uint32_t n, const float* xyr, int32_t bits, uint32_t leaf_size, uint64_t* total) {
    return agglo_collide_host<float, 2>(n, xyr, bits, leaf_size, total);
}
int32_t LBVHa2D_collisions_f64(uint32_t n, const double* xyr, int32_t bits, uint32_t leaf_size, uint64_t* total) {
    return agglo_collide_host<double, 2>(n, xyr, bits, leaf_size, total);
}
int32_t LBVHa3D_collisions_f64(uint32_t n, const double* xyzr, int32_t bits, uint32_t leaf_size, uint64_t* total) {
    return agglo_collide_host<double, 3>(n, xyzr, bits, leaf_size, total);
}

// Device TraversalMinHeap<u32, float, 32> check (include/geometry/
// traversal_stack.hpp:31-90 and tests/geometry/traversal_stack.cpp): push n
// items, pop all; out arrays hold min(n, 32) entries; returns the number of
// pushes accepted in *pushed.
int32_t Stream_heap_order(uint32_t n, const uint32_t* idx, const float* cost, uint32_t* out_idx, float* out_cost,
                          int32_t* pushed) {
    return guarded([&] {
        if (n > 1024) throw std::invalid_argument("at most 1024 items");
        DBuf<uint32_t> di, doi;
        DBuf<float> dc, doc;
        DBuf<int32_t> dp;
        di.resize(n + 1);
        dc.resize(n + 1);
        doi.resize(33);
        doc.resize(33);
        dp.resize(1);
        h2d(di.p, idx, n);
        h2d(dc.p, cost, n);
        hipLaunchKernelGGL(heap_order_kernel, dim3(1), dim3(64), 0, ctx().stream, n, di.p, dc.p, doi.p, doc.p, dp.p);
        const int32_t k = read_scalar(dp.p);
        *pushed = k;
        d2h(out_idx, doi.p, (size_t)k);
        d2h(out_cost, doc.p, (size_t)k);
    });
}

}  // extern "C"
