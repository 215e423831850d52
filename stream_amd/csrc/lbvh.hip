// Karras LBVH kernels and host driver (see lbvh.hpp).
#include <vector>

#include "lbvh.hpp"

namespace stream_amd {

namespace {

template <class S, int D> __device__ __forceinline__ void wave_reduce_box(BoxD<S, D>& b) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
#pragma unroll
        for (int a = 0; a < D; ++a) {
            b.lo[a] = fmin(b.lo[a], __shfl_xor(b.lo[a], o, kWave));
            b.hi[a] = fmax(b.hi[a], __shfl_xor(b.hi[a], o, kWave));
        }
    }
}

// Reduce one box per thread over a 256-thread block; result valid in thread 0.
template <class S, int D> __device__ BoxD<S, D> block_reduce_box(BoxD<S, D> b, BoxD<S, D>* s_tmp) {
    constexpr int NW = 256 / kWave;
    wave_reduce_box<S, D>(b);
    const uint32_t w = threadIdx.x / kWave;
    if (lane_id() == 0) s_tmp[w] = b;
    __syncthreads();
    if (threadIdx.x < kWave) {
        b = threadIdx.x < NW ? s_tmp[threadIdx.x] : box_empty<S, D>();
        wave_reduce_box<S, D>(b);
    }
    __syncthreads();
    return b;
}

template <class S, int D>
__global__ __launch_bounds__(256) void box_partials_kernel(const Sph4<S>* __restrict__ obj, uint32_t n,
                                                           BoxD<S, D>* __restrict__ partials) {
    __shared__ BoxD<S, D> s_tmp[4];
    BoxD<S, D> b = box_empty<S, D>();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        box_combine(b, sphere_box<S, D>(obj[i]));
    b = block_reduce_box<S, D>(b, s_tmp);
    if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

// Morton code of every centre in the frame of the global sphere box
// (lbvh.hpp:185-212) and, when passes > 0, the per-digit histograms of the
// radix sort.  Each block first folds the box partials.
template <class S, int D>
__global__ __launch_bounds__(256) void morton_hist_kernel(const Sph4<S>* __restrict__ obj, uint32_t n,
                                                          const BoxD<S, D>* __restrict__ partials, uint32_t nparts,
                                                          BoxD<S, D>* __restrict__ gbox, int bits,
                                                          uint64_t* __restrict__ codes, uint32_t* __restrict__ idx,
                                                          uint32_t* __restrict__ hist, int passes) {
    __shared__ uint32_t s_h[8 * 256];
    __shared__ BoxD<S, D> s_tmp[4];
    __shared__ BoxD<S, D> s_g;
    for (int i = threadIdx.x; i < 8 * 256; i += blockDim.x) s_h[i] = 0;
    BoxD<S, D> b = box_empty<S, D>();
    for (uint32_t k = threadIdx.x; k < nparts; k += blockDim.x) box_combine(b, partials[k]);
    b = block_reduce_box<S, D>(b, s_tmp);
    if (threadIdx.x == 0) {
        s_g = b;
        if (blockIdx.x == 0) *gbox = b;
    }
    __syncthreads();
    const BoxD<S, D> g = s_g;
    const uint64_t imax = 1ull << bits;
    S f[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const S range = g.hi[a] - g.lo[a];
        S fa = (S)imax / range;
        if (range * fa >= (S)imax) fa = nextafter(fa, (S)0);
        f[a] = fa;
    }
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const Sph4<S> s = obj[i];
        uint32_t q[D];
#pragma unroll
        for (int a = 0; a < D; ++a) q[a] = (uint32_t)((s.c[a] - g.lo[a]) * f[a]);
        uint64_t code;
        if constexpr (D == 2)
            code = spread2(q[0]) | (spread2(q[1]) << 1);
        else
            code = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
        codes[i] = code;
        idx[i] = (uint32_t)i;
        for (int p = 0; p < passes; ++p) atomicAdd(&s_h[p * 256 + (uint32_t)((code >> (8 * p)) & 255u)], 1u);
    }
    __syncthreads();
    if (passes)
        for (int i = threadIdx.x; i < passes * 256; i += blockDim.x)
            if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

template <class S>
__global__ __launch_bounds__(256) void gather_kernel(const Sph4<S>* __restrict__ obj, const uint32_t* __restrict__ perm,
                                                     uint32_t n, Sph4<S>* __restrict__ obj_m) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) obj_m[i] = obj[perm[i]];
}

// Karras 2012 split search (lbvh.hpp:254-307), clz(0) = 64.
__global__ __launch_bounds__(256) void karras_kernel(const uint64_t* __restrict__ m, int n, uint32_t* __restrict__ node,
                                                     uint32_t* __restrict__ leaf_parent,
                                                     uint32_t* __restrict__ internal_parent,
                                                     uint32_t* __restrict__ touched) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n - 1) return;
    touched[i] = 0;
    auto inr = [n](int k) { return k >= 0 && k < n; };
    const uint64_t mi = m[i];
    const int dp1 = (i + 1 >= n) ? -1 : clz64(mi ^ m[i + 1]);
    const int dm1 = (i - 1 < 0) ? -1 : clz64(mi ^ m[i - 1]);
    const int d = (dp1 - dm1 > 0) ? 1 : -1;
    const int dmin = inr(i - d) ? clz64(mi ^ m[i - d]) : -1;
    int lmax = 2;
    while (inr(i + lmax * d) && clz64(mi ^ m[i + lmax * d]) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (inr(i + (l + t) * d) && clz64(mi ^ m[i + (l + t) * d]) > dmin) l += t;
    const int j = i + l * d;
    int gamma = (i + j) >> 1;
    const uint64_t mj = m[j];
    if (mi != mj) {
        const int dnode = clz64(mi ^ mj);
        int s = 0;
        for (int t = (l + 1) / 2; t >= 1; t = (t + 1) / 2) {
            if (inr(i + (s + t) * d) && clz64(mi ^ m[i + (s + t) * d]) > dnode) s += t;
            if (t == 1) break;
        }
        gamma = i + s * d + (d < 0 ? d : 0);
    }
    uint32_t cl = 0;
    const int lo = i < j ? i : j, hi = i > j ? i : j;
    if (gamma == lo) {
        cl |= 1u;
        leaf_parent[gamma] = (uint32_t)i;
    } else {
        internal_parent[gamma] = (uint32_t)i;
    }
    if (gamma + 1 == hi) {
        cl |= 2u;
        leaf_parent[gamma + 1] = (uint32_t)i | 0x80000000u;
    } else {
        internal_parent[gamma + 1] = (uint32_t)i | 0x80000000u;
    }
    if (i == 0) internal_parent[0] = (uint32_t)n;
    node[i] = (uint32_t)gamma | (cl << 30);
}

template <class S, int D, uint32_t MODE> __device__ __forceinline__ BoxD<S, D> leaf_box(const Sph4<S>& s) {
    if constexpr (MODE == kLeafCentre)
        return centre_box<S, D>(s);
    else
        return sphere_box<S, D>(s);
}

// Bottom-up refit (lbvh.hpp:311-351).  The arrival counter is a relaxed
// agent-scope fetch_or; box records travel through coherent_store/load.
template <class S, int D, uint32_t MODE>
__global__ __launch_bounds__(256) void refit_kernel(const Sph4<S>* __restrict__ obj_m, uint32_t n,
                                                    const uint32_t* __restrict__ node,
                                                    const uint32_t* __restrict__ leaf_parent,
                                                    const uint32_t* __restrict__ internal_parent, uint32_t* touched,
                                                    BoxD<S, D>* ibox, uint32_t* flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    BoxD<S, D> data = leaf_box<S, D, MODE>(obj_m[i]);
    uint32_t parent = leaf_parent[i];
    bool right = (parent & 0x80000000u) != 0;
    parent &= 0x7fffffffu;
    // a valid Karras tree is at most 64 + log2(n) deep; the cap only stops
    // a malformed topology (duplicate codes) from cycling
    for (uint32_t level = 0; parent < n; ++level) {
        if (level > 1024u) {
            raise_flag(flags, kErrCapacity);
            break;
        }
        vmem_release_order();
        const uint32_t old = __hip_atomic_fetch_or(&touched[parent], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!old) break;
        vmem_acquire_order();
        const uint32_t w = node[parent];
        const uint32_t sep = w & 0x3fffffffu;
        const uint32_t cl = w >> 30;
        const uint32_t other = sep + (right ? 0u : 1u);
        if (cl & (right ? 1u : 2u))
            box_combine(data, leaf_box<S, D, MODE>(obj_m[other]));
        else
            box_combine(data, coherent_load(&ibox[other]));
        coherent_store(&ibox[parent], data);
        const uint32_t p = internal_parent[parent];
        right = (p & 0x80000000u) != 0;
        parent = p & 0x7fffffffu;
    }
}

}  // namespace

template <class S, int D> void LbvhGpu<S, D>::codes_only(const Sph4<S>* d_obj, uint32_t n_, int bits_, hipStream_t s) {
    n = n_;
    bits = bits_;
    gbox.resize(1);
    partials.resize(kBoxBlocks);
    k0.resize(n ? n : 1);
    v0.resize(n ? n : 1);
    if (!n) return;
    const unsigned nb = grid_for(n, 256, kBoxBlocks);
    hipLaunchKernelGGL((box_partials_kernel<S, D>), dim3(nb), dim3(256), 0, s, d_obj, n, partials.p);
    hipLaunchKernelGGL((morton_hist_kernel<S, D>), dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, d_obj, n,
                       partials.p, nb, gbox.p, bits, k0.p, v0.p, (uint32_t*)nullptr, 0);
}

template <class S, int D>
void LbvhGpu<S, D>::sort_only(const Sph4<S>* d_obj, uint32_t n_, int bits_, hipStream_t s) {
    n = n_;
    bits = bits_;
    if (bits < 1 || bits * D > 64) throw std::invalid_argument("Morton bits per axis out of range");
    obj_m.resize((size_t)n + 1);
    gbox.resize(1);
    partials.resize(kBoxBlocks);
    const uint32_t nn = n ? n : 1;
    k0.resize(nn);
    k1.resize(nn);
    v0.resize(nn);
    v1.resize(nn);
    sortws.resize(radix_workspace_words(nn));
    if (!n) return;
    const int passes = (bits * D + 7) / 8;
    RadixWork w = radix_work_from(sortws.p, n);
    radix_reset(w, passes, s);
    const unsigned nb = grid_for(n, 256, kBoxBlocks);
    hipLaunchKernelGGL((box_partials_kernel<S, D>), dim3(nb), dim3(256), 0, s, d_obj, n, partials.p);
    hipLaunchKernelGGL((morton_hist_kernel<S, D>), dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, d_obj, n,
                       partials.p, nb, gbox.p, bits, k0.p, v0.p, w.hist, passes);
    const bool in1 = radix_sort_pairs(k0.p, v0.p, k1.p, v1.p, n, passes * 8, w, ctx().d_flags, s);
    codes_sorted = in1 ? k1.p : k0.p;
    perm = in1 ? v1.p : v0.p;
    hipLaunchKernelGGL((gather_kernel<S>), dim3(grid_for(n, 256)), dim3(256), 0, s, d_obj, perm, n, obj_m.p);
}

template <class S, int D>
void LbvhGpu<S, D>::build(const Sph4<S>* d_obj, uint32_t n_, int bits_, uint32_t leaf_mode_, uint32_t extra,
                          hipStream_t s) {
    n = n_;
    bits = bits_;
    leaf_mode = leaf_mode_;
    if (bits < 1 || bits * D > 64) throw std::invalid_argument("Morton bits per axis out of range");
    obj_m.resize((size_t)n + extra + 1);
    gbox.resize(1);
    partials.resize(kBoxBlocks);
    const uint32_t nn = n ? n : 1;
    k0.resize(nn);
    k1.resize(nn);
    v0.resize(nn);
    v1.resize(nn);
    sortws.resize(radix_workspace_words(nn));
    node.resize(nn);
    ibox.resize(nn);
    leaf_parent.resize(nn);
    internal_parent.resize(nn);
    touched.resize(nn);
    if (!n) return;
    const int passes = (bits * D + 7) / 8;
    RadixWork w = radix_work_from(sortws.p, n);
    radix_reset(w, passes, s);
    const unsigned nb = grid_for(n, 256, kBoxBlocks);
    section_mark(s, kMarkBuild0);
    hipLaunchKernelGGL((box_partials_kernel<S, D>), dim3(nb), dim3(256), 0, s, d_obj, n, partials.p);
    hipLaunchKernelGGL((morton_hist_kernel<S, D>), dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, d_obj, n,
                       partials.p, nb, gbox.p, bits, k0.p, v0.p, w.hist, passes);
    section_mark(s, kMarkMorton);
    const bool in1 = radix_sort_pairs(k0.p, v0.p, k1.p, v1.p, n, passes * 8, w, ctx().d_flags, s);
    section_mark(s, kMarkSort);
    codes_sorted = in1 ? k1.p : k0.p;
    perm = in1 ? v1.p : v0.p;
    hipLaunchKernelGGL((gather_kernel<S>), dim3(grid_for(n, 256)), dim3(256), 0, s, d_obj, perm, n, obj_m.p);
    section_mark(s, kMarkGather);
    if (n < 2) return;
    hipLaunchKernelGGL(karras_kernel, dim3(grid_for(n - 1, 256)), dim3(256), 0, s, codes_sorted, (int)n, node.p,
                       leaf_parent.p, internal_parent.p, touched.p);
    section_mark(s, kMarkKarras);
    if (leaf_mode == kLeafCentre)
        hipLaunchKernelGGL((refit_kernel<S, D, kLeafCentre>), dim3(grid_for(n, 256)), dim3(256), 0, s, obj_m.p, n,
                           node.p, leaf_parent.p, internal_parent.p, touched.p, ibox.p, ctx().d_flags);
    else
        hipLaunchKernelGGL((refit_kernel<S, D, kLeafSphere>), dim3(grid_for(n, 256)), dim3(256), 0, s, obj_m.p, n,
                           node.p, leaf_parent.p, internal_parent.p, touched.p, ibox.p, ctx().d_flags);
    section_mark(s, kMarkRefit);
}

template struct LbvhGpu<float, 2>;
template struct LbvhGpu<double, 2>;
template struct LbvhGpu<float, 3>;
template struct LbvhGpu<double, 3>;

// ---------------------------------------------------------------------------
// Sphere-sphere collision count (the query of
// tests/geometry/lbvh_simple_collision.cpp:100-140) over a sphere-box LBVH.
// ---------------------------------------------------------------------------
namespace {
template <class S, int D>
__global__ __launch_bounds__(256) void collide_kernel(const Sph4<S>* __restrict__ obj_m, uint32_t n,
                                                     const uint32_t* __restrict__ node,
                                                     const BoxD<S, D>* __restrict__ ibox,
                                                     unsigned long long* __restrict__ total) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c = 0;
    if (tid < n) {
        int stack[64];
        int ss = 0;
        stack[ss++] = 0;
        const Sph4<S> q = obj_m[tid];
        for (uint32_t visits = 0; ss && visits < n; ++visits) {
            const uint32_t cur = (uint32_t)stack[--ss];
            const uint32_t w = node[cur];
            const uint32_t sep = w & 0x3fffffffu, cl = w >> 30;
            for (uint32_t ch = 0; ch < 2; ++ch) {
                if (!(cl & (ch + 1))) {
                    const BoxD<S, D> b = ibox[sep + ch];
                    bool in = true;
                    for (int a = 0; a < D; ++a) {
                        in &= q.c[a] - q.r < b.hi[a];
                        in &= q.c[a] + q.r > b.lo[a];
                    }
                    if (in && ss < 64) stack[ss++] = (int)(sep + ch);
                } else {
                    const Sph4<S> o = obj_m[sep + ch];
                    const S rs = q.r + o.r;
                    S d2 = 0;
                    for (int a = 0; a < D; ++a) {
                        const S dd = q.c[a] - o.c[a];
                        d2 += dd * dd;
                    }
                    if (d2 <= rs * rs && sep + ch != tid) ++c;
                }
            }
        }
    }
    c = wave_sum(c);
    if (lane_id() == 0 && c) atomicAdd(total, (unsigned long long)c);
}

template <class S, int D> std::vector<Sph4<S>> pack_spheres(uint32_t n, const S* xyr) {
    std::vector<Sph4<S>> h(n);
    for (uint32_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) h[i].c[a] = a < D ? xyr[(size_t)(D + 1) * i + a] : (S)0;
        h[i].r = xyr[(size_t)(D + 1) * i + D];
    }
    return h;
}

template <class S, int D>
int32_t lbvh_build_host(uint32_t n, const S* xyr, int32_t bits, int32_t leaf_mode, uint64_t* codes_sorted,
                        uint32_t* perm, int32_t* sep, uint8_t* cl, uint32_t* leaf_parent, uint32_t* internal_parent,
                        S* ibox, S* gbox) {
    return guarded([&] {
        if (n < 2) throw std::invalid_argument("LBVH needs at least 2 objects");
        hipStream_t s = ctx().stream;
        auto h = pack_spheres<S, D>(n, xyr);
        DBuf<Sph4<S>> d_obj;
        d_obj.resize(n);
        h2d(d_obj.p, h.data(), n);
        LbvhGpu<S, D> t;
        t.build(d_obj.p, n, bits, (uint32_t)leaf_mode, 0, s);
        if (codes_sorted) d2h(codes_sorted, t.codes_sorted, n);
        if (perm) d2h(perm, t.perm, n);
        if (leaf_parent) d2h(leaf_parent, t.leaf_parent.p, n);
        if (internal_parent) d2h(internal_parent, t.internal_parent.p, n - 1);
        if (sep || cl) {
            std::vector<uint32_t> w(n - 1);
            d2h(w.data(), t.node.p, n - 1);
            for (uint32_t i = 0; i + 1 < n; ++i) {
                if (sep) sep[i] = (int32_t)(w[i] & 0x3fffffffu);
                if (cl) cl[i] = (uint8_t)(w[i] >> 30);
            }
        }
        if (ibox) d2h((BoxD<S, D>*)ibox, t.ibox.p, n - 1);
        if (gbox) d2h((BoxD<S, D>*)gbox, t.gbox.p, 1);
        if (read_scalar(ctx().d_flags + kFlagError) != 0) throw std::runtime_error("device error flag raised");
    });
}

template <class S, int D> int32_t lbvh_morton_host(uint32_t n, const S* xyr, int32_t bits, uint64_t* codes) {
    return guarded([&] {
        hipStream_t s = ctx().stream;
        auto h = pack_spheres<S, D>(n, xyr);
        DBuf<Sph4<S>> d_obj;
        d_obj.resize(n ? n : 1);
        h2d(d_obj.p, h.data(), n);
        LbvhGpu<S, D> t;
        t.codes_only(d_obj.p, n, bits, s);
        d2h(codes, t.k0.p, n);
    });
}

template <class S, int D> int32_t lbvh_collide_host(uint32_t n, const S* xyr, int32_t bits, uint64_t* total) {
    return guarded([&] {
        if (n < 2) throw std::invalid_argument("LBVH needs at least 2 objects");
        hipStream_t s = ctx().stream;
        auto h = pack_spheres<S, D>(n, xyr);
        DBuf<Sph4<S>> d_obj;
        d_obj.resize(n);
        h2d(d_obj.p, h.data(), n);
        LbvhGpu<S, D> t;
        t.build(d_obj.p, n, bits, kLeafSphere, 0, s);
        DBuf<unsigned long long> tot;
        tot.resize(1);
        fill<unsigned long long>(tot.p, 1, 0ull, s);
        hipLaunchKernelGGL((collide_kernel<S, D>), dim3(grid_for(n, 256)), dim3(256), 0, s, t.obj_m.p, n, t.node.p,
                           t.ibox.p, tot.p);
        *total = read_scalar(tot.p);
    });
}
}  // namespace

}  // namespace stream_amd

using namespace stream_amd;

extern "C" {

// LBVH over n spheres given as host rows (x, y, r).  Outputs (host, any may
// be null): Morton codes in sorted order, the sort permutation (map_sorted),
// the Karras topology (internal_sep, child_is_leaf, leaf/internal parents
// with bit 31 = right child) and the refitted internal boxes as rows
// (xmin, ymin, xmax, ymax), plus the global sphere box.  leaf_mode 0 boxes
// the centres (AlphaShape2D::InternalNodeDataFunctor), 1 the spheres
// (ComputeBB_Functor).  No reference C symbol exists for the LBVH
// (src/geometry/lbvh.cpp:22-27 only instantiates templates); this is the
// test surface of include/geometry/lbvh.hpp:354-360.
int32_t LBVH2D_build_f32(uint32_t n, const float* xyr, int32_t bits, int32_t leaf_mode, uint64_t* codes_sorted,
                         uint32_t* perm, int32_t* sep, uint8_t* cl, uint32_t* lp, uint32_t* ip, float* ibox,
                         float* gbox) {
    return lbvh_build_host<float, 2>(n, xyr, bits, leaf_mode, codes_sorted, perm, sep, cl, lp, ip, ibox, gbox);
}
int32_t LBVH2D_build_f64(uint32_t n, const double* xyr, int32_t bits, int32_t leaf_mode, uint64_t* codes_sorted,
                         uint32_t* perm, int32_t* sep, uint8_t* cl, uint32_t* lp, uint32_t* ip, double* ibox,
                         double* gbox) {
    return lbvh_build_host<double, 2>(n, xyr, bits, leaf_mode, codes_sorted, perm, sep, cl, lp, ip, ibox, gbox);
}
int32_t LBVH3D_build_f64(uint32_t n, const double* xyzr, int32_t bits, int32_t leaf_mode, uint64_t* codes_sorted,
                         uint32_t* perm, int32_t* sep, uint8_t* cl, uint32_t* lp, uint32_t* ip, double* ibox,
                         double* gbox) {
    return lbvh_build_host<double, 3>(n, xyzr, bits, leaf_mode, codes_sorted, perm, sep, cl, lp, ip, ibox, gbox);
}
// Unsorted Morton codes (lbvh.hpp:178-213) of n host spheres.
int32_t LBVH2D_morton_f32(uint32_t n, const float* xyr, int32_t bits, uint64_t* codes) {
    return lbvh_morton_host<float, 2>(n, xyr, bits, codes);
}
int32_t LBVH2D_morton_f64(uint32_t n, const double* xyr, int32_t bits, uint64_t* codes) {
    return lbvh_morton_host<double, 2>(n, xyr, bits, codes);
}
int32_t LBVH3D_morton_f64(uint32_t n, const double* xyzr, int32_t bits, uint64_t* codes) {
    return lbvh_morton_host<double, 3>(n, xyzr, bits, codes);
}
// Total sphere-sphere collisions (self excluded) found through the LBVH.
int32_t LBVH2D_collisions_f32(uint32_t n, const float* xyr, int32_t bits, uint64_t* total) {
    return lbvh_collide_host<float, 2>(n, xyr, bits, total);
}
int32_t LBVH2D_collisions_f64(uint32_t n, const double* xyr, int32_t bits, uint64_t* total) {
    return lbvh_collide_host<double, 2>(n, xyr, bits, total);
}

}  // extern "C"
