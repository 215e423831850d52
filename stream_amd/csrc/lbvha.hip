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
//
// TRBVH (include/geometry/trbvh.hpp:64-681) is the same topology with
// surface-area costs and subtree sizes, then treelet restructuring rounds.
#include <vector>

#include "lbvha.hpp"
#include "traversal.hpp"

namespace stream_amd {

namespace {

__device__ __forceinline__ void st_u32(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_u32(const uint32_t* p) {
    return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_f32(float* p, float v) { st_u32(reinterpret_cast<uint32_t*>(p), __float_as_uint(v)); }
__device__ __forceinline__ float ld_f32(const float* p) {
    return __uint_as_float(ld_u32(reinterpret_cast<const uint32_t*>(p)));
}

// TRBVH cost constants (trbvh.hpp:72-75).
constexpr float kSahCi = 1.2f;
constexpr float kSahCt = 1.0f;
constexpr int kTreelet = 5;
constexpr uint32_t kTreeletAll = (1u << kTreelet) - 1;

// (c * ext0) * ext1: the reference's area term, 2D even in 3D.
template <class S, int D> __device__ __forceinline__ float sah_term(float c, const BoxD<S, D>& b) {
    return (c * (float)(b.hi[0] - b.lo[0])) * (float)(b.hi[1] - b.lo[1]);
}

template <class S, int D>
__global__ __launch_bounds__(256) void agglo_init_kernel(const Sph4<S>* __restrict__ obj_m,
                                                         const uint64_t* __restrict__ codes, uint32_t n,
                                                         uint32_t leaf_mode, BoxD<S, D>* __restrict__ data,
                                                         uint64_t* __restrict__ delta, uint32_t* __restrict__ touched,
                                                         float* __restrict__ sah, uint32_t* __restrict__ subtree) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    data[i] = leaf_mode == kLeafCentre ? centre_box<S, D>(obj_m[i]) : sphere_box<S, D>(obj_m[i]);
    if (sah) {
        sah[i] = 0.0f;
        subtree[i] = 1u;
    }
    if (i + 1 < n) {
        delta[i] = codes[i] ^ codes[i + 1];
        touched[i] = 0;
    }
}

template <class S, int D>
__global__ __launch_bounds__(256) void agglo_build_kernel(uint32_t n, const uint64_t* __restrict__ delta,
                                                          uint32_t* children, uint32_t* range, uint32_t* parent_of,
                                                          uint32_t* split, uint32_t* touched, BoxD<S, D>* data,
                                                          uint32_t* root, float* sah, uint32_t* subtree,
                                                          uint32_t* flags) {
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
        if (sah) {
            // TRBVH cost (trbvh.hpp:321-343): the thread arriving second
            // carries its own running cost (Ct * area of its object at a
            // leaf) and adds Ci * area(parent) + stored cost of the sibling
            // (0 at a leaf).  The reference lets either child arrive second;
            // here the right child is always the carrier, which is what a
            // sequential run of the reference's loop produces.
            const uint32_t lc = side == 0 ? idx : sibling;
            const uint32_t rc = side == 0 ? sibling : idx;
            const float carry = rc < n ? sah_term<S, D>(kSahCt, coherent_load(&data[rc])) : ld_f32(&sah[rc]);
            const float stored = lc < n ? 0.0f : ld_f32(&sah[lc]);
            st_f32(&sah[parent], carry + (sah_term<S, D>(kSahCi, box) + stored));
            st_u32(&subtree[parent], ld_u32(&subtree[lc]) + ld_u32(&subtree[rc]));
        }
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

// ---------------------------------------------------------------------------
// TRBVH treelet restructuring (trbvh.hpp:390-633).  One thread per leaf
// climbs as in the build; the second thread to reach a node owns it and,
// when the node has at least min_size objects below it, optimises the
// treelet of 5 leaves rooted there:
//   * formation: start from the node's two children and expand, three
//     times, the costliest treelet leaf that is an internal node (the
//     reference takes the costliest of all and would read an object leaf's
//     missing children when every internal cost is zero, :432-447);
//   * dynamic programme over the 31 leaf subsets (Karras & Aila 2013):
//     c(s) = min(Ci A(s) + min_p (c(p) + c(s^p)), Ct A(s) N(s)), partitions
//     p over the non-empty subsets of s minus its lowest member in
//     increasing order, first minimum kept; A = ext0 * ext1;
//   * rebuild from the partitions: p goes right, s^p left, the treelet's
//     internal nodes reused in stack order, their boxes, sizes and costs
//     recomputed (:540-602).
// Each thread keeps its c/p tables in its own LDS column (row = subset).
// A treelet only touches the owned node's subtree, which is final, so the
// result does not depend on the interleaving.
// ---------------------------------------------------------------------------
constexpr uint32_t kTreeletBlock = 64;

template <int D>
__device__ void trbvh_treelet(uint32_t idx, uint32_t n, uint32_t* children, uint32_t* parent_of,
                              BoxD<float, D>* data, float* sah, uint32_t* subtree, float (*s_copt)[kTreeletBlock],
                              uint8_t (*s_popt)[kTreeletBlock], uint32_t* flags) {
    const uint32_t t = threadIdx.x;
    uint32_t inner[kTreelet - 1];
    uint32_t leaf[kTreelet];
    float cost[kTreelet];
    inner[0] = idx;
    leaf[0] = ld_u32(&children[2 * (idx - n)]);
    leaf[1] = ld_u32(&children[2 * (idx - n) + 1]);
    cost[0] = ld_f32(&sah[leaf[0]]);
    cost[1] = ld_f32(&sah[leaf[1]]);
#pragma unroll
    for (int m = 1; m < kTreelet - 1; ++m) {
        int best = -1;
        float bc = 0.0f;
        uint32_t b = 0;
#pragma unroll
        for (int l = 0; l <= m; ++l)
            if (leaf[l] >= n && (best < 0 || cost[l] > bc)) {
                best = l;
                bc = cost[l];
                b = leaf[l];
            }
        if (best < 0) {  // fewer objects than treelet leaves below idx
            raise_flag(flags, kErrCapacity);
            return;
        }
        inner[m] = b;
#pragma unroll
        for (int l = 0; l < m; ++l)
            if (l == best) {
                leaf[l] = leaf[m];
                cost[l] = cost[m];
            }
        leaf[m] = ld_u32(&children[2 * (b - n)]);
        leaf[m + 1] = ld_u32(&children[2 * (b - n) + 1]);
        cost[m] = ld_f32(&sah[leaf[m]]);
        cost[m + 1] = ld_f32(&sah[leaf[m + 1]]);
    }
    BoxD<float, D> lb[kTreelet];
    uint32_t ls[kTreelet];
#pragma unroll
    for (int l = 0; l < kTreelet; ++l) {
        lb[l] = coherent_load(&data[leaf[l]]);
        ls[l] = ld_u32(&subtree[leaf[l]]);
        s_copt[1u << l][t] = cost[l];
    }
    auto subset_box = [&](uint32_t part, uint32_t& size) {
        BoxD<float, D> b = box_empty<float, D>();
        size = 0;
#pragma unroll
        for (int l = 0; l < kTreelet; ++l)
            if (part & (1u << l)) {
                box_combine(b, lb[l]);
                size += ls[l];
            }
        return b;
    };
    for (int k = 2; k <= kTreelet; ++k) {
        for (uint32_t s = 3; s <= kTreeletAll; ++s) {
            if (__popc(s) != k) continue;
            float cs = FLT_MAX;
            uint32_t ps = 0;
            const uint32_t delta = (s - 1) & s;
            for (uint32_t p = (0u - delta) & s; p; p = (p - delta) & s) {
                const float c = s_copt[p][t] + s_copt[s ^ p][t];
                if (c < cs) {
                    cs = c;
                    ps = p;
                }
            }
            uint32_t size;
            const BoxD<float, D> b = subset_box(s, size);
            const float area = (b.hi[0] - b.lo[0]) * (b.hi[1] - b.lo[1]);
            const float split = kSahCi * area + cs;
            const float flat = kSahCt * area * (float)size;
            s_copt[s][t] = fminf(split, flat);
            s_popt[s][t] = (uint8_t)ps;
        }
    }
    st_f32(&sah[idx], s_copt[kTreeletAll][t]);
    uint32_t st_part[kTreelet], st_right[kTreelet], st_parent[kTreelet];
    int ss = 0, used = 0;
    const uint32_t p0 = s_popt[kTreeletAll][t];
    st_part[0] = p0;
    st_right[0] = 1;
    st_parent[0] = idx;
    st_part[1] = kTreeletAll & ~p0;
    st_right[1] = 0;
    st_parent[1] = idx;
    ss = 2;
    while (ss > 0) {
        --ss;
        const uint32_t part = st_part[ss], right = st_right[ss], par = st_parent[ss];
        const bool is_leaf = __popc(part) == 1;
        uint32_t node = 0;
        if (is_leaf) {
            const int l = __ffs(part) - 1;
#pragma unroll
            for (int q = 0; q < kTreelet; ++q)
                if (q == l) node = leaf[q];
        } else {
            ++used;
#pragma unroll
            for (int q = 1; q < kTreelet - 1; ++q)
                if (q == used) node = inner[q];
        }
        st_u32(&children[2 * (par - n) + right], node);
        st_u32(&parent_of[node], par);
        if (!is_leaf) {
            uint32_t size;
            const BoxD<float, D> b = subset_box(part, size);
            coherent_store(&data[node], b);
            st_u32(&subtree[node], size);
            st_f32(&sah[node], s_copt[part][t]);
            const uint32_t pl = s_popt[part][t];
            if (ss + 2 > kTreelet) {
                raise_flag(flags, kErrCapacity);
                return;
            }
            st_part[ss] = pl;
            st_right[ss] = 1;
            st_parent[ss] = node;
            ++ss;
            st_part[ss] = part & ~pl;
            st_right[ss] = 0;
            st_parent[ss] = node;
            ++ss;
        }
    }
}

template <int D>
__global__ __launch_bounds__(kTreeletBlock) void trbvh_round_kernel(uint32_t n, uint32_t min_size,
                                                                   uint32_t* children, uint32_t* parent_of,
                                                                   BoxD<float, D>* data, float* sah,
                                                                   uint32_t* subtree, uint32_t* touched,
                                                                   uint32_t* flags) {
    __shared__ float s_copt[kTreeletAll + 1][kTreeletBlock];
    __shared__ uint8_t s_popt[kTreeletAll + 1][kTreeletBlock];
    const uint32_t i = blockIdx.x * kTreeletBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t idx = i;
    uint32_t parent = ld_u32(&parent_of[i]);
    uint32_t sub = ld_u32(&subtree[i]);
    for (uint32_t level = 0;; ++level) {
        if (level > 4096u) {
            raise_flag(flags, kErrCapacity);
            break;
        }
        if (sub >= min_size && idx >= n)
            trbvh_treelet<D>(idx, n, children, parent_of, data, sah, subtree, s_copt, s_popt, flags);
        if (parent == 2u * n) break;
        vmem_release_order();
        const uint32_t old =
            __hip_atomic_fetch_add(&touched[parent - n], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!old) break;
        vmem_acquire_order();
        idx = parent;
        sub = ld_u32(&subtree[idx]);
        parent = ld_u32(&parent_of[idx]);
    }
}

template <int D> struct Trbvh {
    Agglo<float, D> t;
    DBuf<float> sah;         // [2n-1]
    DBuf<uint32_t> subtree;  // [2n-1]
    float ms_build = 0.0f, ms_rounds = 0.0f;

    // Build (trbvh.hpp:637-679): agglomerative hierarchy with costs, then
    // `rounds` restructuring passes with min_size 5, 10, 20, ... (the
    // reference runs 3).
    void build(const Sph4<float>* d_obj, uint32_t n, int rounds, hipStream_t s) {
        if (n < 2) throw std::invalid_argument("TRBVH needs at least 2 objects");
        if (rounds < 0 || rounds > 16) throw std::invalid_argument("rounds must be in [0, 16]");
        sah.resize(2ull * n);
        subtree.resize(2ull * n);
        hipEvent_t e[3];
        for (auto& x : e) SA_HIP(hipEventCreate(&x));
        SA_HIP(hipEventRecord(e[0], s));
        t.build(d_obj, n, 64 / D, kLeafSphere, s, sah.p, subtree.p);
        SA_HIP(hipEventRecord(e[1], s));
        uint32_t min_size = kTreelet;
        for (int r = 0; r < rounds; ++r, min_size *= 2) {
            fill<uint32_t>(t.touched.p, n - 1, 0u, s);
            hipLaunchKernelGGL((trbvh_round_kernel<D>), dim3(grid_for(n, kTreeletBlock)), dim3(kTreeletBlock), 0, s,
                               n, min_size, t.children.p, t.parent_of.p, t.data.p, sah.p, subtree.p, t.touched.p,
                               ctx().d_flags);
        }
        SA_HIP(hipEventRecord(e[2], s));
        SA_HIP(hipEventSynchronize(e[2]));
        SA_HIP(hipEventElapsedTime(&ms_build, e[0], e[1]));
        SA_HIP(hipEventElapsedTime(&ms_rounds, e[1], e[2]));
        for (auto& x : e) SA_HIP(hipEventDestroy(x));
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

template <int D>
int32_t trbvh_build_host(uint32_t n, const float* xyr, int32_t rounds, uint32_t* perm, uint32_t* children,
                         uint32_t* range, uint32_t* parent, uint32_t* root, float* boxes, float* sah,
                         uint32_t* subtree, double* ms) {
    return guarded([&] {
        hipStream_t s = ctx().stream;
        auto h = pack_rows<float, D>(n, xyr);
        DBuf<Sph4<float>> d_obj;
        d_obj.resize(n ? n : 1);
        h2d(d_obj.p, h.data(), n);
        Trbvh<D> T;
        T.build(d_obj.p, n, rounds, s);
        check_device_flags();
        if (perm) d2h(perm, T.t.base.perm, n);
        if (children) d2h(children, T.t.children.p, 2ull * (n - 1));
        if (range) d2h(range, T.t.range.p, 2ull * (n - 1));
        if (parent) d2h(parent, T.t.parent_of.p, 2ull * n - 1);
        if (root) d2h(root, T.t.root.p, 1);
        if (boxes) d2h((BoxD<float, D>*)boxes, T.t.data.p, 2ull * n - 1);
        if (sah) d2h(sah, T.sah.p, 2ull * n - 1);
        if (subtree) d2h(subtree, T.subtree.p, 2ull * n - 1);
        if (ms) {
            ms[0] = T.ms_build;
            ms[1] = T.ms_rounds;
        }
    });
}

template <int D> int32_t trbvh_collide_host(uint32_t n, const float* xyr, int32_t rounds, uint64_t* total) {
    return guarded([&] {
        hipStream_t s = ctx().stream;
        auto h = pack_rows<float, D>(n, xyr);
        DBuf<Sph4<float>> d_obj;
        d_obj.resize(n ? n : 1);
        h2d(d_obj.p, h.data(), n);
        Trbvh<D> T;
        T.build(d_obj.p, n, rounds, s);
        DBuf<unsigned long long> tot;
        tot.resize(1);
        fill<unsigned long long>(tot.p, 1, 0ull, s);
        // Leaf buckets of one object: the build's ranges are not kept up to
        // date by the restructuring (as in the reference) and are not used.
        hipLaunchKernelGGL((agglo_collide_kernel<float, D>), dim3(grid_for(n, 256)), dim3(256), 0, s,
                           T.t.base.obj_m.p, n, T.t.children.p, T.t.range.p, T.t.data.p, T.t.root.p, 1u, tot.p,
                           ctx().d_flags);
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

template <class S, int D>
void Agglo<S, D>::build(const Sph4<S>* d_obj, uint32_t n, int bits, uint32_t leaf_mode, hipStream_t s, float* sah,
                        uint32_t* subtree) {
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
                       base.codes_sorted, n, leaf_mode, data.p, delta.p, touched.p, sah, subtree);
    hipLaunchKernelGGL((agglo_build_kernel<S, D>), dim3(grid_for(n, 256)), dim3(256), 0, s, n, delta.p, children.p,
                       range.p, parent_of.p, split.p, touched.p, data.p, root.p, sah, subtree, ctx().d_flags);
}
template struct Agglo<float, 2>;
template struct Agglo<double, 2>;
template struct Agglo<float, 3>;
template struct Agglo<double, 3>;

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

// TRBVH over n host spheres in float, rows (x, y[, z], r): the
// agglomerative build with surface-area costs, then `rounds` treelet
// restructuring passes (the reference runs 3).  Outputs as LBVHa2D_build_*
// (range as built, not updated by the restructuring) plus the node costs
// sah [2n-1] and subtree object counts [2n-1]; ms (may be null) gets the
// GPU time of the build and of the rounds.  Test surface of
// include/geometry/trbvh.hpp:64-681 (Point2DTRBVH / Sphere2DTRBVH /
// Sphere3DTRBVH), which exports no C symbol.
int32_t TRBVH2D_build_f32(uint32_t n, const float* xyr, int32_t rounds, uint32_t* perm, uint32_t* children,
                          uint32_t* range, uint32_t* parent, uint32_t* root, float* boxes, float* sah,
                          uint32_t* subtree, double* ms) {
    return trbvh_build_host<2>(n, xyr, rounds, perm, children, range, parent, root, boxes, sah, subtree, ms);
}
int32_t TRBVH3D_build_f32(uint32_t n, const float* xyzr, int32_t rounds, uint32_t* perm, uint32_t* children,
                          uint32_t* range, uint32_t* parent, uint32_t* root, float* boxes, float* sah,
                          uint32_t* subtree, double* ms) {
    return trbvh_build_host<3>(n, xyzr, rounds, perm, children, range, parent, root, boxes, sah, subtree, ms);
}
// Sphere-sphere collision count (self excluded) over a TRBVH
// (tests/geometry/trbvh_random_collision.cpp:87-139).
int32_t TRBVH2D_collisions_f32(uint32_t n, const float* xyr, int32_t rounds, uint64_t* total) {
    return trbvh_collide_host<2>(n, xyr, rounds, total);
}
int32_t TRBVH3D_collisions_f32(uint32_t n, const float* xyzr, int32_t rounds, uint64_t* total) {
    return trbvh_collide_host<3>(n, xyzr, rounds, total);
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
