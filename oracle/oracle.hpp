// CPU oracle for Stream's PFEM mesh-and-solve path.
//
// TEST INFRASTRUCTURE ONLY.  This file is a plain C++ restatement of the
// reference algorithms, templated on the scalar type (float reproduces the
// reference's fp_tt=float build, double serves the FP64 workloads).  It is
// compiled into oracle/lib/liboracle.so and used by tests/, by the smoke check
// and by bench.py's cpu_baseline leg only.  Nothing in stream_amd/ includes,
// links or calls it.
//
// Each function cites the reference file:line it restates.  How each part is
// pinned (reference golden vectors, the compiled reference predicates in
// oracle/_ref, closed forms, brute force) is listed next to it.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace orc {

// --------------------------------------------------------------------------
// Scalar constants.
// eps = 2^-p (unit roundoff, include/geometry/predicates.hpp:27 for float).
// two_delta = fl(2*sqrt((1+4eps)/(1-3eps))) -- the search-radius inflation of
// src/mesh/DirectAlphaShape2D.cpp:188 (float literal there); for double it
// is 2 + 7eps rounded to nearest = 2 + 2^-50.
// --------------------------------------------------------------------------
template <class S> struct Consts;
template <> struct Consts<float> {
    static constexpr float eps = 5.9604644775390625e-08f;                // 2^-24
    static constexpr float two_delta = 2.0000004172325445139859201780136934301654f;
};
template <> struct Consts<double> {
    static constexpr double eps = 1.1102230246251565404236316680908203125e-16;  // 2^-53
    static constexpr double two_delta = 2.0 + 0x1p-50;
};

// --------------------------------------------------------------------------
// Contracted evaluation of the reference's inexact arithmetic.  The reference
// is built by nvcc, whose default (--fmad=true) may fuse a product that
// feeds an addition or subtraction into one fma (one rounding instead of
// two).  Its certified predicate signs do not depend on that, but three
// inexact decisions do: the candidate test |c_i - c_j|^2 < (2 delta a)^2
// (DirectAlphaShape2D.cpp:212-213, DirectAlphaShape3D.cpp:204-205), the
// box-distance test of the BVH walk (2D.cpp:201-206, 3D.cpp:192-199) and the
// circumradius filters (2D.cpp:554-563, 3D.cpp:611-645).  `mode` bits choose
// which of these sites are evaluated fused; kFirstPair / kSubRight choose
// which product of a two-product sum or difference is the fused one.  Mode 0
// is the unfused statement-order evaluation.
// --------------------------------------------------------------------------
namespace fmac {
enum : int { kLeaf = 1, kBox = 2, kFilter = 4, kFirstPair = 8, kSubRight = 16 };

// x0*x0 + x1*x1 + ... + x(k-1)*x(k-1), summed left to right
template <class S> inline S sumsq(int mode, bool site, const S* x, int k) {
    if (!site) {
        S s = x[0] * x[0];
        for (int i = 1; i < k; ++i) s = s + x[i] * x[i];
        return s;
    }
    S s = (mode & kFirstPair) ? std::fma(x[0], x[0], x[1] * x[1]) : std::fma(x[1], x[1], x[0] * x[0]);
    for (int i = 2; i < k; ++i) s = std::fma(x[i], x[i], s);
    return s;
}
// a*b - c*d
template <class S> inline S diff(int mode, bool site, S a, S b, S c, S d) {
    if (!site) return a * b - c * d;
    return (mode & kSubRight) ? std::fma(-c, d, a * b) : std::fma(a, b, -(c * d));
}
// a*b + c*d
template <class S> inline S dot2(int mode, bool site, S a, S b, S c, S d) {
    if (!site) return a * b + c * d;
    return (mode & kFirstPair) ? std::fma(a, b, c * d) : std::fma(c, d, a * b);
}
// s + a*b
template <class S> inline S acc(bool site, S s, S a, S b) { return site ? std::fma(a, b, s) : s + a * b; }
}  // namespace fmac

// --------------------------------------------------------------------------
// Morton codes (include/geometry/morton.hpp:27-92).  Written here as a plain
// bit loop: bit k of axis a lands at bit k*dim + a.  Pinned against the
// reference's magic-bit encoders compiled in oracle/_ref.
// --------------------------------------------------------------------------
inline uint64_t interleave(const uint32_t* q, int dim, int bits) {
    uint64_t code = 0;
    for (int k = 0; k < bits; ++k)
        for (int a = 0; a < dim; ++a)
            code |= (uint64_t)((q[a] >> k) & 1u) << (k * dim + a);
    return code;
}

inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

// --------------------------------------------------------------------------
// Objects: spheres (centre c, radius r).  get_bbox = [c-r, c+r]
// (include/geometry/primitives.hpp:136-141).
// --------------------------------------------------------------------------
template <class S, int D> struct Sphere {
    S c[D];
    S r;
};
template <class S, int D> struct Box {
    S lo[D], hi[D];
    static Box empty() {
        Box b;
        for (int a = 0; a < D; ++a) {
            b.lo[a] = std::numeric_limits<S>::max();
            b.hi[a] = -std::numeric_limits<S>::max();
        }
        return b;
    }
    // combineBox (include/geometry/bbox.hpp:88-96)
    void combine(const Box& o) {
        for (int a = 0; a < D; ++a) lo[a] = std::fmin(lo[a], o.lo[a]);
        for (int a = 0; a < D; ++a) hi[a] = std::fmax(hi[a], o.hi[a]);
    }
};
template <class S, int D> Box<S, D> sphere_box(const Sphere<S, D>& s) {
    Box<S, D> b;
    for (int a = 0; a < D; ++a) {
        b.lo[a] = s.c[a] - s.r;
        b.hi[a] = s.c[a] + s.r;
    }
    return b;
}
template <class S, int D> Box<S, D> centre_box(const Sphere<S, D>& s) {
    Box<S, D> b;
    for (int a = 0; a < D; ++a) b.lo[a] = b.hi[a] = s.c[a];
    return b;
}

// --------------------------------------------------------------------------
// Karras LBVH (include/geometry/lbvh.hpp:141-360).
// --------------------------------------------------------------------------
template <class S, int D> struct Lbvh {
    uint32_t n = 0;
    Box<S, D> global;                        // box of all sphere boxes (lbvh.hpp:141-175)
    std::vector<uint64_t> codes;             // unsorted (lbvh.hpp:178-213)
    std::vector<uint64_t> codes_sorted;      // (lbvh.hpp:216-235)
    std::vector<uint32_t> perm;              // map_sorted
    std::vector<Sphere<S, D>> obj_m;         // objects in Morton order (lbvh.hpp:240-242)
    std::vector<int32_t> sep;                // internal_sep (lbvh.hpp:287)
    std::vector<uint8_t> child_is_leaf;      // bit0 left, bit1 right
    std::vector<uint32_t> leaf_parent;       // bit 31 = right child
    std::vector<uint32_t> internal_parent;   // bit 31 = right child; [0] = n
    std::vector<Box<S, D>> internal_box;     // refit data (lbvh.hpp:311-351)
};

// Global bbox of the sphere boxes (lbvh.hpp:141-175).
template <class S, int D> Box<S, D> global_box(const std::vector<Sphere<S, D>>& obj) {
    Box<S, D> g = Box<S, D>::empty();
    for (const auto& s : obj) g.combine(sphere_box(s));
    return g;
}

// Quantisation + interleave (lbvh.hpp:185-212).  bits = bits per axis; the
// reference hard-codes bits = 64/dim.
template <class S, int D>
void morton_codes(const std::vector<Sphere<S, D>>& obj, const Box<S, D>& g, int bits,
                  std::vector<uint64_t>& out) {
    const uint64_t imax = 1ULL << bits;
    S f[D];
    for (int a = 0; a < D; ++a) {
        const S range = g.hi[a] - g.lo[a];
        S fa = (S)imax / range;
        if (range * fa >= (S)imax) fa = std::nextafter(fa, (S)0);
        f[a] = fa;
    }
    out.resize(obj.size());
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < obj.size(); ++i) {
        uint32_t q[D];
        for (int a = 0; a < D; ++a) {
            const S c = (obj[i].c[a] - g.lo[a]) * f[a];
            q[a] = (uint32_t)c;
        }
        out[i] = interleave(q, D, bits);
    }
}

// Karras 2012 hierarchy (lbvh.hpp:246-308) with clz(0) defined as 64.
template <class S, int D> void build_hierarchy(Lbvh<S, D>& t) {
    const int n = (int)t.n;
    const auto& m = t.codes_sorted;
    t.sep.assign(n - 1, 0);
    t.child_is_leaf.assign(n - 1, 0);
    t.leaf_parent.assign(n, 0);
    t.internal_parent.assign(n - 1, 0);
    auto inr = [n](int k) { return k >= 0 && k < n; };
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n - 1; ++i) {
        const uint64_t mi = m[i];
        const int dp1 = (i + 1 >= n) ? -1 : clz64(mi ^ m[i + 1]);
        const int dm1 = (i - 1 < 0) ? -1 : clz64(mi ^ m[i - 1]);
        const int d = (dp1 - dm1 > 0) ? 1 : -1;
        const int dmin = inr(i - d) ? clz64(mi ^ m[i - d]) : -1;
        int lmax = 2;
        while (inr(i + lmax * d) && clz64(mi ^ m[i + lmax * d]) > dmin) lmax *= 2;
        int l = 0;
        for (int tt = lmax / 2; tt >= 1; tt /= 2)
            if (inr(i + (l + tt) * d) && clz64(mi ^ m[i + (l + tt) * d]) > dmin) l += tt;
        const int j = i + l * d;
        int gamma = (i + j) >> 1;
        const uint64_t mj = m[j];
        if (mi != mj) {
            const int dnode = clz64(mi ^ mj);
            int s = 0;
            for (int tt = (l + 1) / 2; tt >= 1; tt = (tt + 1) / 2) {
                if (inr(i + (s + tt) * d) && clz64(mi ^ m[i + (s + tt) * d]) > dnode) s += tt;
                if (tt == 1) break;
            }
            gamma = i + s * d + (d < 0 ? d : 0);
        }
        t.sep[i] = gamma;
        uint8_t cl = 0;
        if (gamma == std::min(i, j)) {
            cl |= 1;
            t.leaf_parent[gamma] = (uint32_t)i;
        } else {
            t.internal_parent[gamma] = (uint32_t)i;
        }
        if (gamma + 1 == std::max(i, j)) {
            cl |= 2;
            t.leaf_parent[gamma + 1] = (uint32_t)i | 0x80000000u;
        } else {
            t.internal_parent[gamma + 1] = (uint32_t)i | 0x80000000u;
        }
        if (i == 0) t.internal_parent[0] = (uint32_t)n;
        t.child_is_leaf[i] = cl;
    }
}

// Bottom-up refit (lbvh.hpp:311-351), executed sequentially leaf after leaf:
// the first arrival at a node stops, the second combines with its sibling.
// leaf_box selects the data functor: centre box for the alpha-shape
// (DirectAlphaShape2D.hpp:87-99), sphere box for ComputeBB_Functor.
template <class S, int D, class LeafBox> void refit(Lbvh<S, D>& t, LeafBox leaf_box) {
    const uint32_t n = t.n;
    std::vector<uint8_t> touched(n - 1, 0);
    t.internal_box.assign(n - 1, Box<S, D>::empty());
    for (uint32_t i = 0; i < n; ++i) {
        Box<S, D> data = leaf_box(t.obj_m[i]);
        uint32_t parent = t.leaf_parent[i];
        bool is_right = (parent & 0x80000000u) != 0;
        parent &= 0x7fffffffu;
        while (parent < n) {
            const uint8_t old = touched[parent];
            touched[parent] = 1;
            if (!old) break;
            const uint32_t child = (uint32_t)t.sep[parent];
            const uint32_t other = child + (is_right ? 0u : 1u);
            if (t.child_is_leaf[parent] & (is_right ? 1 : 2))
                data.combine(leaf_box(t.obj_m[other]));
            else
                data.combine(t.internal_box[other]);
            t.internal_box[parent] = data;
            parent = t.internal_parent[parent];
            is_right = (parent & 0x80000000u) != 0;
            parent &= 0x7fffffffu;
        }
    }
}

// Full build (lbvh.hpp:354-360): bbox, codes, stable sort, gather, hierarchy,
// refit.  n >= 2 is required.
template <class S, int D, class LeafBox>
void build_lbvh(Lbvh<S, D>& t, const std::vector<Sphere<S, D>>& obj, int bits, LeafBox leaf_box) {
    t.n = (uint32_t)obj.size();
    t.global = global_box(obj);
    morton_codes(obj, t.global, bits, t.codes);
    t.perm.resize(t.n);
    for (uint32_t i = 0; i < t.n; ++i) t.perm[i] = i;
    std::stable_sort(t.perm.begin(), t.perm.end(),
                     [&](uint32_t a, uint32_t b) { return t.codes[a] < t.codes[b]; });
    t.codes_sorted.resize(t.n);
    t.obj_m.resize(t.n);
    for (uint32_t i = 0; i < t.n; ++i) {
        t.codes_sorted[i] = t.codes[t.perm[i]];
        t.obj_m[i] = obj[t.perm[i]];
    }
    build_hierarchy(t);
    refit(t, leaf_box);
}

// --------------------------------------------------------------------------
// Agglomerative LBVH (include/geometry/lbvh_agglo.hpp:62-361).  The leaf
// climbs are run one after the other in leaf order; the topology depends
// only on the deltas and the boxes only on min/max, so this gives the same
// arrays as any parallel interleaving.
// --------------------------------------------------------------------------
template <class S, int D> struct LbvhA {
    uint32_t n = 0;
    Box<S, D> global;
    std::vector<uint64_t> codes_sorted;
    std::vector<uint32_t> perm;
    std::vector<Sphere<S, D>> obj_m;
    std::vector<uint32_t> children, range, parent, split;  // [n-1][2], [n-1][2], [2n-1], [n-1]
    std::vector<Box<S, D>> data;                            // [2n-1]
    uint32_t root = 0;
};

template <class S, int D, class LeafBox>
void build_lbvha(LbvhA<S, D>& t, const std::vector<Sphere<S, D>>& obj, int bits, LeafBox leaf_box) {
    const uint32_t n = (uint32_t)obj.size();
    t.n = n;
    t.global = global_box(obj);
    std::vector<uint64_t> codes;
    morton_codes(obj, t.global, bits, codes);
    t.perm.resize(n);
    for (uint32_t i = 0; i < n; ++i) t.perm[i] = i;
    std::stable_sort(t.perm.begin(), t.perm.end(), [&](uint32_t a, uint32_t b) { return codes[a] < codes[b]; });
    t.codes_sorted.resize(n);
    t.obj_m.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        t.codes_sorted[i] = codes[t.perm[i]];
        t.obj_m[i] = obj[t.perm[i]];
    }
    t.children.assign(2ull * (n - 1), 0);
    t.range.assign(2ull * (n - 1), 0);
    t.parent.assign(2ull * n - 1, 0);
    t.split.assign(n - 1, 0);
    t.data.assign(2ull * n - 1, Box<S, D>::empty());
    std::vector<uint64_t> delta(n - 1);
    for (uint32_t i = 0; i + 1 < n; ++i) delta[i] = t.codes_sorted[i] ^ t.codes_sorted[i + 1];
    std::vector<uint32_t> touched(n - 1, 0);
    for (uint32_t i = 0; i < n; ++i) t.data[i] = leaf_box(t.obj_m[i]);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t idx = i, imin = i, imax = i;
        Box<S, D> data = t.data[i];
        while (true) {
            if (imin == 0 && imax == n - 1) {
                t.parent[idx] = 2 * n;
                t.root = idx;
                break;
            }
            const bool right = imin == 0 || (imax != n - 1 && delta[imax] < delta[imin - 1]);
            const uint32_t parent = right ? imax + n : imin + n - 1;
            const uint32_t k = parent - n;
            t.range[2 * k + (right ? 0 : 1)] = right ? imin : imax;
            t.children[2 * k + (right ? 0 : 1)] = idx;
            t.parent[idx] = parent;
            if (touched[k]++) {
                const uint32_t sib = t.children[2 * k + (right ? 1 : 0)];
                data.combine(t.data[sib]);
                t.data[parent] = data;
                t.split[k] = right ? imax : imin - 1;
                idx = parent;
                imin = t.range[2 * k];
                imax = t.range[2 * k + 1];
            } else {
                break;
            }
        }
    }
}

// Collision count with leaf buckets (tests/geometry/
// lbvh_agglo_simple_collision.cpp:105-160), self excluded.
template <class S, int D> uint64_t lbvha_collisions(const LbvhA<S, D>& t, uint32_t leaf_size) {
    const uint32_t n = t.n;
    uint64_t total = 0;
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 256)
    for (uint32_t tid = 0; tid < n; ++tid) {
        uint32_t stack[64];
        uint32_t ss = 0;
        stack[ss++] = t.root;
        const Sphere<S, D>& q = t.obj_m[tid];
        while (ss) {
            const uint32_t cur = stack[--ss];
            if (cur < n) continue;
            for (int ch = 0; ch < 2; ++ch) {
                const uint32_t cid = t.children[2 * (cur - n) + ch];
                uint32_t rmin = cid, rmax = cid;
                if (cid >= n) {
                    rmin = t.range[2 * (cid - n)];
                    rmax = t.range[2 * (cid - n) + 1];
                }
                const bool leaf = cid < n || (rmax - rmin < leaf_size);
                if (!leaf) {
                    const Box<S, D>& b = t.data[cid];
                    bool in = true;
                    for (int a = 0; a < D; ++a) {
                        in &= q.c[a] - q.r < b.hi[a];
                        in &= q.c[a] + q.r > b.lo[a];
                    }
                    if (in) stack[ss++] = cid;
                } else {
                    for (uint32_t o = rmin; o <= rmax; ++o) {
                        const Sphere<S, D>& s = t.obj_m[o];
                        const S rs = q.r + s.r;
                        S d2 = 0;
                        for (int a = 0; a < D; ++a) {
                            const S dd = q.c[a] - s.c[a];
                            d2 += dd * dd;
                        }
                        if (d2 <= rs * rs && o != tid) ++total;
                    }
                }
            }
        }
    }
    return total;
}

// --------------------------------------------------------------------------
// TRBVH (include/geometry/trbvh.hpp:64-681): the agglomerative topology with
// surface-area costs and subtree sizes (:294-388), then treelet
// restructuring rounds (:390-633) with minimum subtree sizes 5, 10, 20
// (:670-674).  fp_tt = float, the only build of the reference.
//
// Every parallel loop of the reference is run here leaf after leaf.  Two
// facts make that a faithful restatement:
//  * the cost sum at a node depends on which child's thread arrives second
//    (that thread's running cost counts its own leaf's area, the sibling's
//    stored cost does not, :340-343 vs :321-323).  In leaf order the second
//    arrival always comes from the right child; a parallel run of the
//    reference may pick either, this restatement (and the GPU build) fixes
//    the right child;
//  * a treelet rooted at a node is formed only after both of the node's
//    children are final, and it touches that node's subtree only, so the
//    rounds do not depend on the interleaving.
// One guard departs from the reference: the leaf chosen for expansion is the
// costliest treelet leaf among internal nodes.  The reference picks the
// costliest of all treelet leaves (:432-439) and would index an object leaf
// as an internal node when every internal cost is zero (boxes of zero
// area); whenever some internal leaf carries the maximum the two agree.
// --------------------------------------------------------------------------
constexpr float kTrbvhCi = 1.2f;  // trbvh.hpp:72
constexpr float kTrbvhCt = 1.0f;  // trbvh.hpp:73
constexpr int kTreeletLeaves = 5; // trbvh.hpp:74

template <int D> struct Trbvh {
    LbvhA<float, D> t;              // topology, ranges (as built), parents, root, boxes
    std::vector<float> sah;         // [2n-1]
    std::vector<uint32_t> subtree;  // [2n-1]
};

// Area used by the reference in every dimension: ext[0] * ext[1].
template <int D> float trbvh_area(const Box<float, D>& b) {
    const float e0 = b.hi[0] - b.lo[0];
    const float e1 = b.hi[1] - b.lo[1];
    return e0 * e1;
}

// Costs and subtree sizes of the freshly built hierarchy (trbvh.hpp:294-388):
// leaf cost 0 and size 1 (:283-284, :323); a thread's running cost starts
// at Ct * area(own object) (:321) and gains Ci * area(node) + cost(sibling)
// at every node it completes (:340-343).
template <int D> void trbvh_costs(Trbvh<D>& T) {
    const uint32_t n = T.t.n;
    T.sah.assign(2ull * n - 1, 0.0f);
    T.subtree.assign(2ull * n - 1, 1u);
    std::vector<uint32_t> touched(n - 1, 0);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t idx = i, sub = 1;
        const Box<float, D>& bl = T.t.data[i];
        float run = (kTrbvhCt * (bl.hi[0] - bl.lo[0])) * (bl.hi[1] - bl.lo[1]);
        while (idx != T.t.root) {
            const uint32_t p = T.t.parent[idx];
            const uint32_t k = p - n;
            if (touched[k]++ == 0) break;
            const uint32_t l = T.t.children[2 * k], r = T.t.children[2 * k + 1];
            const uint32_t sib = l == idx ? r : l;
            sub += T.subtree[sib];
            const Box<float, D>& b = T.t.data[p];
            const float node = (kTrbvhCi * (b.hi[0] - b.lo[0])) * (b.hi[1] - b.lo[1]);
            run = run + (node + T.sah[sib]);
            T.sah[p] = run;
            T.subtree[p] = sub;
            idx = p;
        }
    }
}

// Optimal partition of a treelet of 5 leaves (trbvh.hpp:460-533): surface
// areas of the 31 leaf subsets, then for every subset s of two or more
// leaves c(s) = min(Ci A(s) + min over splits (c(p) + c(s^p)), Ct A(s) N(s))
// with N(s) the number of objects under s.  copt/popt/size get 32 entries
// (index = leaf subset bit mask).
template <int D>
void treelet_dp(const Box<float, D>* lbox, const float* cost, const uint32_t* lsize, float* copt, uint32_t* popt,
                uint32_t* size) {
    constexpr uint32_t kAll = (1u << kTreeletLeaves) - 1;
    float area[kAll + 1] = {};
    for (uint32_t s = 0; s <= kAll; ++s) {
        copt[s] = 0.0f;
        popt[s] = 0;
        size[s] = 0;
    }
    for (uint32_t s = 1; s <= kAll; ++s) {
        Box<float, D> b = Box<float, D>::empty();
        for (int l = 0; l < kTreeletLeaves; ++l)
            if (s & (1u << l)) {
                b.combine(lbox[l]);
                size[s] += lsize[l];
            }
        area[s] = trbvh_area<D>(b);
    }
    for (int l = 0; l < kTreeletLeaves; ++l) copt[1u << l] = cost[l];
    // Karras & Aila 2013 (HPG), partition search: the lowest member of s
    // stays on one side, p runs over the non-empty subsets of the others in
    // increasing order; ties keep the first.
    // Subsets are visited by increasing size; within one size the order is
    // immaterial (each depends on smaller subsets only).
    for (int k = 2; k <= kTreeletLeaves; ++k) {
        for (uint32_t s = 3; s <= kAll; ++s) {
            if (__builtin_popcount(s) != k) continue;
            float cs = std::numeric_limits<float>::max();
            uint32_t ps = 0;
            const uint32_t delta = (s - 1) & s;
            for (uint32_t p = (0u - delta) & s; p; p = (p - delta) & s) {
                const float c = copt[p] + copt[s ^ p];
                if (c < cs) {
                    cs = c;
                    ps = p;
                }
            }
            const float split = kTrbvhCi * area[s] + cs;
            const float flat = kTrbvhCt * area[s] * (float)size[s];
            copt[s] = std::fmin(split, flat);
            popt[s] = ps;
        }
    }
}

// One treelet at internal node idx (trbvh.hpp:408-605).
template <int D> void trbvh_treelet(Trbvh<D>& T, uint32_t idx) {
    const uint32_t n = T.t.n;
    auto child = [&](uint32_t node, int side) { return T.t.children[2 * (node - n) + side]; };
    uint32_t inner[kTreeletLeaves - 1] = {};
    uint32_t leaf[kTreeletLeaves] = {};
    float cost[kTreeletLeaves] = {};
    inner[0] = idx;
    leaf[0] = child(idx, 0);
    leaf[1] = child(idx, 1);
    cost[0] = T.sah[leaf[0]];
    cost[1] = T.sah[leaf[1]];
    for (int m = 1; m < kTreeletLeaves - 1; ++m) {
        int best = -1;
        for (int l = 0; l <= m; ++l)
            if (leaf[l] >= n && (best < 0 || cost[l] > cost[best])) best = l;
        const uint32_t b = leaf[best];
        inner[m] = b;
        leaf[best] = leaf[m];
        cost[best] = cost[m];
        leaf[m] = child(b, 0);
        cost[m] = T.sah[leaf[m]];
        leaf[m + 1] = child(b, 1);
        cost[m + 1] = T.sah[leaf[m + 1]];
    }
    constexpr uint32_t kAll = (1u << kTreeletLeaves) - 1;
    Box<float, D> lbox[kTreeletLeaves];
    uint32_t lsize[kTreeletLeaves];
    for (int l = 0; l < kTreeletLeaves; ++l) {
        lbox[l] = T.t.data[leaf[l]];
        lsize[l] = T.subtree[leaf[l]];
    }
    float copt[kAll + 1];
    uint32_t popt[kAll + 1], size[kAll + 1];
    treelet_dp<D>(lbox, cost, lsize, copt, popt, size);
    // Rebuild from the partitions (:540-602): the partition p_opt goes to
    // the right child, its complement to the left; internal nodes are
    // reused in the order they come off the stack.
    T.sah[idx] = copt[kAll];
    struct Item {
        uint32_t part, right, parent;
    } st[kTreeletLeaves];
    int ss = 0, used = 0;
    st[ss++] = {popt[kAll], 1u, idx};
    st[ss++] = {kAll & ~popt[kAll], 0u, idx};
    while (ss) {
        const Item it = st[--ss];
        const bool is_leaf = __builtin_popcount(it.part) == 1;
        const uint32_t node = is_leaf ? leaf[__builtin_ctz(it.part)] : inner[++used];
        T.t.children[2 * (it.parent - n) + it.right] = node;
        T.t.parent[node] = it.parent;
        if (!is_leaf) {
            Box<float, D> b = Box<float, D>::empty();
            for (int l = 0; l < kTreeletLeaves; ++l)
                if (it.part & (1u << l)) b.combine(T.t.data[leaf[l]]);
            T.t.data[node] = b;
            T.subtree[node] = size[it.part];
            T.sah[node] = copt[it.part];
            const uint32_t pl = popt[it.part];
            st[ss++] = {pl, 1u, node};
            st[ss++] = {it.part & ~pl, 0u, node};
        }
    }
}

// One restructuring round (trbvh.hpp:390-633), leaf after leaf.
template <int D> void trbvh_round(Trbvh<D>& T, uint32_t min_size) {
    const uint32_t n = T.t.n;
    std::vector<uint32_t> touched(n - 1, 0);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t idx = i, parent = T.t.parent[i], sub = T.subtree[i];
        while (true) {
            if (sub >= min_size) trbvh_treelet<D>(T, idx);
            if (parent == 2 * n) break;
            if (touched[parent - n]++ == 0) break;
            idx = parent;
            sub = T.subtree[idx];
            parent = T.t.parent[idx];
        }
    }
}

// Full build (trbvh.hpp:637-679); rounds = 3 in the reference.
template <int D> void build_trbvh(Trbvh<D>& T, const std::vector<Sphere<float, D>>& obj, int rounds) {
    build_lbvha(T.t, obj, 64 / D, [](const Sphere<float, D>& s) { return sphere_box(s); });
    trbvh_costs<D>(T);
    uint32_t min_size = kTreeletLeaves;
    for (int r = 0; r < rounds; ++r, min_size *= 2) trbvh_round<D>(T, min_size);
}

// Collision count over a TRBVH (tests/geometry/trbvh_random_collision.cpp:
// 87-139): leaves are objects, self excluded.
template <int D> uint64_t trbvh_collisions(const Trbvh<D>& T) {
    const uint32_t n = T.t.n;
    uint64_t total = 0;
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 256)
    for (uint32_t tid = 0; tid < n; ++tid) {
        std::vector<uint32_t> stack;
        stack.push_back(T.t.root);
        const Sphere<float, D>& q = T.t.obj_m[tid];
        while (!stack.empty()) {
            const uint32_t cur = stack.back();
            stack.pop_back();
            if (cur < n) continue;
            for (int ch = 0; ch < 2; ++ch) {
                const uint32_t cid = T.t.children[2 * (cur - n) + ch];
                if (cid >= n) {
                    const Box<float, D>& b = T.t.data[cid];
                    bool in = true;
                    for (int a = 0; a < D; ++a) {
                        in &= q.c[a] - q.r < b.hi[a];
                        in &= q.c[a] + q.r > b.lo[a];
                    }
                    if (in) stack.push_back(cid);
                } else {
                    const Sphere<float, D>& s = T.t.obj_m[cid];
                    const float rs = q.r + s.r;
                    float d2 = 0;
                    for (int a = 0; a < D; ++a) {
                        const float dd = q.c[a] - s.c[a];
                        d2 += dd * dd;
                    }
                    if (d2 <= rs * rs && cid != tid) ++total;
                }
            }
        }
    }
    return total;
}

// --------------------------------------------------------------------------
// Robust predicates (include/geometry/predicates.hpp:342-583).
//
// Filters are restated operation for operation (including the fma form of
// the incircle filter, predicates.hpp:556-582).  The exact stage is written
// differently from the reference on purpose: the determinant is expanded
// into its monomials, every monomial is formed exactly as a sum of doubles
// (error-free products via fma), and all components are accumulated with a
// zero-eliminating grow-expansion.  Only the sign of an exact stage is ever
// used by the callers, so exact stages return -1, 0 or +1.  Pinned against
// the compiled reference predicates (oracle/_ref) and exact rationals.
// --------------------------------------------------------------------------
namespace exact {
inline void two_sum(double a, double b, double& x, double& y) {
    x = a + b;
    const double bv = x - a;
    const double av = x - bv;
    y = (a - av) + (b - bv);
}
// Add scalar b into the nonoverlapping expansion h[0..n) in place.
inline int grow(int n, double* h, double b) {
    double q = b;
    int k = 0;
    for (int i = 0; i < n; ++i) {
        double hh;
        two_sum(q, h[i], q, hh);
        if (hh != 0.0) h[k++] = hh;
    }
    if (q != 0.0 || k == 0) h[k++] = q;
    return k;
}
// Exact product of the given factors, emitted as components into acc.
struct Acc {
    std::vector<double> h;
    int n = 0;
    Acc() { h.resize(64); }
    void add(double v) {
        if (v == 0.0) return;
        if ((int)h.size() < n + 2) h.resize(h.size() * 2);
        n = grow(n, h.data(), v);
    }
    // add sign * prod(f[0..k)) exactly
    void add_product(const double* f, int k, double sign) {
        double parts[64];
        int np = 1;
        parts[0] = sign * f[0];
        for (int j = 1; j < k; ++j) {
            double next[64];
            int nn = 0;
            for (int p = 0; p < np; ++p) {
                const double x = parts[p] * f[j];
                const double y = std::fma(parts[p], f[j], -x);
                next[nn++] = x;
                if (y != 0.0) next[nn++] = y;
            }
            np = nn;
            std::memcpy(parts, next, sizeof(double) * nn);
        }
        for (int p = 0; p < np; ++p) add(parts[p]);
    }
    int sign() const {
        for (int i = n - 1; i >= 0; --i)
            if (h[i] != 0.0) return h[i] > 0 ? 1 : -1;
        return 0;
    }
};

// det | ax ay 1 ; bx by 1 ; cx cy 1 |
template <class S> int orient2d_sign(const S* a, const S* b, const S* c) {
    Acc acc;
    const double ax = a[0], ay = a[1], bx = b[0], by = b[1], cx = c[0], cy = c[1];
    const double t[6][2] = {{ax, by}, {bx, cy}, {cx, ay}, {ax, cy}, {bx, ay}, {cx, by}};
    for (int k = 0; k < 6; ++k) acc.add_product(t[k], 2, k < 3 ? 1.0 : -1.0);
    return acc.sign();
}

// det | px py px^2+py^2 1 | for p = a,b,c,d  (sign convention of Shewchuk's
// incircle: positive when d lies inside the circle through a,b,c taken
// counter-clockwise).  Laplace expansion along the lift column.
template <class S> int incircle_sign(const S* a, const S* b, const S* c, const S* d) {
    const double P[4][2] = {{a[0], a[1]}, {b[0], b[1]}, {c[0], c[1]}, {d[0], d[1]}};
    Acc acc;
    // incircle(a,b,c,d) = det4(rows p: x, y, lift, 1); Laplace expansion on
    // the lift column: det4 = sum_k (-1)^k lift_k M_k with
    // M_k = det3(x, y, 1 of the other three rows).
    for (int k = 0; k < 4; ++k) {
        int r[3], m = 0;
        for (int q = 0; q < 4; ++q)
            if (q != k) r[m++] = q;
        // M_k = x0 y1 + x1 y2 + x2 y0 - x0 y2 - x1 y0 - x2 y1  (rows r0,r1,r2)
        const double mx[6][2] = {{P[r[0]][0], P[r[1]][1]}, {P[r[1]][0], P[r[2]][1]},
                                 {P[r[2]][0], P[r[0]][1]}, {P[r[0]][0], P[r[2]][1]},
                                 {P[r[1]][0], P[r[0]][1]}, {P[r[2]][0], P[r[1]][1]}};
        const double cof = ((k % 2) == 0) ? 1.0 : -1.0;
        for (int t = 0; t < 6; ++t) {
            const double ms = t < 3 ? 1.0 : -1.0;
            for (int comp = 0; comp < 2; ++comp) {
                const double f[4] = {P[k][comp], P[k][comp], mx[t][0], mx[t][1]};
                acc.add_product(f, 4, cof * ms);
            }
        }
    }
    return acc.sign();
}
}  // namespace exact

template <class S> struct PredConst {
    static constexpr S eps = Consts<S>::eps;
    static constexpr S ccwerrboundA = (S)((3.0 + 16.0 * (double)eps) * (double)eps);
    static constexpr S iccerrboundA = (S)((10.0 + 96.0 * (double)eps) * (double)eps);
};

// orient2d filter + exact (predicates.hpp:385-418).  Positive for a,b,c
// counter-clockwise.  The exact branch returns the sign only.
template <class S> S orient2d(const S* pa, const S* pb, const S* pc) {
    const S detleft = (pa[0] - pc[0]) * (pb[1] - pc[1]);
    const S detright = (pa[1] - pc[1]) * (pb[0] - pc[0]);
    const S det = detleft - detright;
    S detsum;
    if (detleft > (S)0) {
        if (detright <= (S)0) return det;
        detsum = detleft + detright;
    } else if (detleft < (S)0) {
        if (detright >= (S)0) return det;
        detsum = -detleft - detright;
    } else {
        return det;
    }
    const S errbound = PredConst<S>::ccwerrboundA * detsum;
    if ((det >= errbound) || (-det >= errbound)) return det;
    return (S)exact::orient2d_sign(pa, pb, pc);
}

// incircle filter (predicates.hpp:544-583): returns det, writes err.
template <class S> S incircle_filter(const S* pa, const S* pb, const S* pc, const S* pd, S* err) {
    const S adx = pa[0] - pd[0], bdx = pb[0] - pd[0], cdx = pc[0] - pd[0];
    const S ady = pa[1] - pd[1], bdy = pb[1] - pd[1], cdy = pc[1] - pd[1];
    const S bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const S alift = std::fma(adx, adx, ady * ady);
    const S cdxady = cdx * ady, adxcdy = adx * cdy;
    const S blift = std::fma(bdx, bdx, bdy * bdy);
    const S adxbdy = adx * bdy, bdxady = bdx * ady;
    const S clift = std::fma(cdx, cdx, cdy * cdy);
    const S det = std::fma(alift, (bdxcdy - cdxbdy), std::fma(blift, (cdxady - adxcdy), clift * (adxbdy - bdxady)));
    const S perm = std::fma(alift, (std::fabs(bdxcdy) + std::fabs(cdxbdy)),
                            std::fma(blift, (std::fabs(cdxady) + std::fabs(adxcdy)),
                                     (std::fabs(adxbdy) + std::fabs(bdxady)) * clift));
    *err = PredConst<S>::iccerrboundA * perm;
    return det;
}

// --------------------------------------------------------------------------
// 2D alpha-shape (src/mesh/DirectAlphaShape2D.cpp:121-689).
// --------------------------------------------------------------------------
struct Edge8 {
    uint8_t a, b;
};

template <class S> struct AlphaShape2D {
    static constexpr uint32_t n_inf = 3;
    static constexpr uint32_t max_fill = 250;  // DirectAlphaShape2D.cpp:351
    uint32_t n = 0;
    int morton_bits = 32;
    int contract = 0;  // fmac mode of the inexact tests
    Lbvh<S, 2> lbvh;
    std::vector<Sphere<S, 2>> coords_m;     // Morton order + 3 infinity nodes
    std::vector<uint32_t> count;            // count pass (cpp:178-225)
    std::vector<uint32_t> nineig;           // 3 + filled candidates (cpp:356)
    std::vector<std::vector<uint32_t>> cand;   // candidate ids in traversal order
    std::vector<std::vector<Edge8>> tri;       // kept local triangles after filter
    std::vector<std::vector<uint8_t>> active;  // per-candidate edge counts
    std::vector<uint8_t> node_is_bnd, node_nelem, node_nelem_out;
    std::vector<uint32_t> node_nfneig;
    std::vector<uint32_t> row, elemrow, neig;
    std::vector<uint8_t> edge_is_bnd;
    std::vector<uint32_t> elem;  // 3 per triangle
    uint32_t n_elems = 0, n_edges = 0;
    uint64_t exact_calls = 0;

    // BVH traversal shared by the count and fill passes (cpp:178-225, 310-357).
    template <class Visit> void traverse(uint32_t tid, Visit visit, bool cap) const {
        int stack[64];
        int ss = 0;
        stack[ss++] = 0;
        const Sphere<S, 2> si = coords_m[tid];
        const S two_delta = Consts<S>::two_delta;
        const S rs = two_delta * si.r;
        uint32_t nloc = n_inf;
        while (ss != 0) {
            const uint32_t cur = (uint32_t)stack[--ss];
            const uint32_t sepv = (uint32_t)lbvh.sep[cur];
            const uint8_t cl = lbvh.child_is_leaf[cur];
            for (uint32_t ch = 0; ch < 2; ++ch) {
                if (!(cl & (ch + 1))) {
                    const Box<S, 2>& bb = lbvh.internal_box[sepv + ch];
                    const S dmaxx = std::fmax(si.c[0] - bb.hi[0], (S)0);
                    const S dmaxy = std::fmax(si.c[1] - bb.hi[1], (S)0);
                    const S dminx = std::fmax(bb.lo[0] - si.c[0], (S)0);
                    const S dminy = std::fmax(bb.lo[1] - si.c[1], (S)0);
                    const S dd[4] = {dmaxx, dminx, dmaxy, dminy};
                    const bool hit = fmac::sumsq(contract, contract & fmac::kBox, dd, 4) < rs * rs;
                    stack[ss] = (int)(sepv + ch);
                    ss += hit;
                } else if (tid != sepv + ch) {
                    const Sphere<S, 2>& sj = coords_m[sepv + ch];
                    const S ma = std::fmin(si.r, sj.r);
                    const S dxy[2] = {si.c[0] - sj.c[0], si.c[1] - sj.c[1]};
                    const S sq = fmac::sumsq(contract, contract & fmac::kLeaf, dxy, 2);
                    if (sq < (two_delta * ma) * (two_delta * ma)) {
                        visit(sepv + ch);
                        ++nloc;
                    }
                }
            }
            if (cap && nloc > max_fill) break;
        }
    }

    void init(const std::vector<Sphere<S, 2>>& nodes) {
        n = (uint32_t)nodes.size();
        build_lbvh(lbvh, nodes, morton_bits, [](const Sphere<S, 2>& s) { return centre_box(s); });
        coords_m = lbvh.obj_m;
        // infinity nodes (cpp:144-163): avg and range of the sphere bbox
        const Box<S, 2>& g = lbvh.global;
        S avg[2], range[2];
        for (int a = 0; a < 2; ++a) {
            avg[a] = (g.hi[a] + g.lo[a]) * (S)0.5;
            range[a] = g.hi[a] - g.lo[a];
        }
        const S mr = std::fmax(range[0], range[1]);
        Sphere<S, 2> i0{{avg[0] - 2 * mr, avg[1] - 2 * mr}, (S)0};
        Sphere<S, 2> i1{{avg[0] + 2 * mr, avg[1] - 2 * mr}, (S)0};
        Sphere<S, 2> i2{{avg[0], avg[1] + 2 * mr}, (S)0};
        coords_m.push_back(i0);
        coords_m.push_back(i1);
        coords_m.push_back(i2);
        count.assign(n, 0);
#pragma omp parallel for schedule(dynamic, 1024)
        for (uint32_t tid = 0; tid < n; ++tid) {
            uint32_t c = 0;
            traverse(tid, [&](uint32_t) { ++c; }, false);
            count[tid] = c;
        }
    }

    const S* pt(uint32_t g) const { return coords_m[g].c; }
    uint32_t neig_of(uint32_t tid, uint32_t j) const { return j >= n_inf ? cand[tid][j - n_inf] : n + j; }

    // Local Bowyer-Watson insertion with the simplified SoS, then the alpha
    // filter (cpp:373-600).
    void local_delaunay(uint32_t si, uint64_t& nexact) {
        std::vector<Edge8> T;
        T.reserve(cand[si].size() + 8);
        T.push_back({1, 2});
        T.push_back({2, 0});
        T.push_back({0, 1});
        active[si].assign(cand[si].size(), 0);
        const uint32_t nn = nineig[si];
        if (nn - n_inf < 2) {
            node_is_bnd[si] = 1;
            node_nelem_out[si] = 0;
            node_nelem[si] = 0;
            node_nfneig[si] = 0;
            tri[si].clear();
            return;
        }
        uint64_t tag[4];
        for (uint32_t ni = n_inf; ni < nn; ++ni) {
            tag[0] = tag[1] = tag[2] = tag[3] = 0;
            const uint32_t sj = neig_of(si, ni);
            uint32_t rsize = 0;
            size_t ti = 0;
            while (ti < T.size()) {
                const Edge8 e = T[ti];
                uint32_t i1 = si, i2 = neig_of(si, e.a), i3 = neig_of(si, e.b), i4 = sj;
                const S* pa = pt(i1);
                const S* pb = pt(i2);
                const S* pc = pt(i3);
                const S* pd = pt(i4);
                const S orientation = orient2d(pa, pb, pc);
                S err;
                S det = incircle_filter(pa, pb, pc, pd, &err);
                if (!((det > err) || (-det > err))) {
                    ++nexact;
                    bool swap = false;
                    auto cs = [&](uint32_t& x, uint32_t& y, const S*& px, const S*& py) {
                        if (x > y) {
                            std::swap(x, y);
                            std::swap(px, py);
                            swap = !swap;
                        }
                    };
                    cs(i1, i3, pa, pc);
                    cs(i2, i4, pb, pd);
                    cs(i1, i2, pa, pb);
                    cs(i3, i4, pc, pd);
                    cs(i2, i3, pb, pc);
                    det = (S)exact::incircle_sign(pa, pb, pc, pd);
                    int depth = 0;
                    while (det == (S)0) {
                        ++depth;
                        if (depth == 1)
                            det = orient2d(pb, pc, pd);
                        else if (depth == 2)
                            det = -orient2d(pa, pc, pd);
                        else
                            det = ((int)swap + (int)(orientation < (S)0)) == 1 ? (S)1 : (S)-1;
                    }
                    if (swap) det = -det;
                }
                if (orientation < (S)0) det = -det;
                if (det > (S)0) {
                    tag[e.a >> 6] ^= 1ULL << (e.a & 63);
                    tag[e.b >> 6] ^= 1ULL << (e.b & 63);
                    ++rsize;
                    T[ti] = T.back();
                    T.pop_back();
                } else {
                    ++ti;
                }
            }
            if (rsize) {
                for (uint32_t j = 0; j < 4; ++j) {
                    const uint64_t v = tag[j];
                    const uint32_t end = std::min((j + 1) * 64, ni);
                    if (!v) continue;
                    for (uint32_t r = j * 64; r < end; ++r)
                        if (v & (1ULL << (r & 63))) T.push_back({(uint8_t)r, (uint8_t)ni});
                }
            }
        }
        // filter (cpp:517-594)
        tag[0] = tag[1] = tag[2] = tag[3] = 0;
        bool bnd = false;
        uint32_t nout = 0, nloc = 0;
        size_t ti = 0;
        while (ti < T.size()) {
            const Edge8 e = T[ti];
            if (e.a < n_inf || e.b < n_inf) {
                bnd = true;
                T[ti] = T.back();
                T.pop_back();
                continue;
            }
            uint32_t i1 = si, i2 = neig_of(si, e.a), i3 = neig_of(si, e.b);
            bool swap = false;
            if (i3 < i2) { std::swap(i2, i3); swap = !swap; }
            if (i2 < i1) { std::swap(i1, i2); swap = !swap; }
            if (i3 < i2) { std::swap(i2, i3); swap = !swap; }
            const Sphere<S, 2>& p1 = coords_m[i1];
            const Sphere<S, 2>& p2 = coords_m[i2];
            const Sphere<S, 2>& p3 = coords_m[i3];
            const S ux = p3.c[0] - p2.c[0], uy = p3.c[1] - p2.c[1];
            const S vx = p1.c[0] - p2.c[0], vy = p1.c[1] - p2.c[1];
            const bool fz = contract & fmac::kFilter;
            S det = fmac::diff(contract, fz, ux, vy, uy, vx);
            if (swap) det = -det;
            const S d12[2] = {p1.c[0] - p2.c[0], p1.c[1] - p2.c[1]};
            const S d23[2] = {p2.c[0] - p3.c[0], p2.c[1] - p3.c[1]};
            const S d31[2] = {p3.c[0] - p1.c[0], p3.c[1] - p1.c[1]};
            const S n12 = fmac::sumsq(contract, fz, d12, 2);
            const S n23 = fmac::sumsq(contract, fz, d23, 2);
            const S n31 = fmac::sumsq(contract, fz, d31, 2);
            const S rsq = ((n12 * n23) * n31) / (((S)4 * det) * det);
            const S aeff = std::fmin(p3.r, std::fmin(p1.r, p2.r));
            if (rsq <= aeff * aeff) {
                tag[e.a >> 6] |= 1ULL << (e.a & 63);
                tag[e.b >> 6] |= 1ULL << (e.b & 63);
                active[si][e.a - n_inf] += 1;
                active[si][e.b - n_inf] += 1;
                ++nloc;
                if (det < (S)0) T[ti] = {e.b, e.a};
                ++ti;
                nout += (si == i1);
            } else {
                bnd = true;
                T[ti] = T.back();
                T.pop_back();
            }
        }
        node_is_bnd[si] = bnd;
        node_nelem_out[si] = (uint8_t)nout;
        node_nelem[si] = (uint8_t)nloc;
        node_nfneig[si] = (uint32_t)(__builtin_popcountll(tag[0]) + __builtin_popcountll(tag[1]) +
                                     __builtin_popcountll(tag[2]) + __builtin_popcountll(tag[3]));
        tri[si] = T;
    }

    void compute() {
        cand.assign(n, {});
        nineig.assign(n, 0);
#pragma omp parallel for schedule(dynamic, 1024)
        for (uint32_t tid = 0; tid < n; ++tid) {
            auto& c = cand[tid];
            c.reserve(count[tid]);
            traverse(tid, [&](uint32_t g) { c.push_back(g); }, true);
            nineig[tid] = n_inf + (uint32_t)c.size();
        }
        tri.assign(n, {});
        active.assign(n, {});
        node_is_bnd.assign(n, 0);
        node_nelem.assign(n, 0);
        node_nelem_out.assign(n, 0);
        node_nfneig.assign(n, 0);
        uint64_t nex = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : nex)
        for (uint32_t si = 0; si < n; ++si) local_delaunay(si, nex);
        exact_calls = nex;
        // scans (cpp:604-636)
        elemrow.assign(n + 1, 0);
        row.assign(n + 1, 0);
        for (uint32_t i = 0; i < n; ++i) {
            elemrow[i + 1] = elemrow[i] + node_nelem_out[i];
            row[i + 1] = row[i] + node_nfneig[i];
        }
        n_elems = elemrow[n];
        n_edges = row[n];
        compress();
    }

    // ELL -> CSR (cpp:642-689)
    void compress() {
        neig.assign(n_edges, 0);
        edge_is_bnd.assign(n_edges, 0);
        elem.assign(3ull * n_elems, 0);
#pragma omp parallel for schedule(dynamic, 1024)
        for (uint32_t si = 0; si < n; ++si) {
            uint32_t put = row[si];
            for (uint32_t i = n_inf; i < nineig[si]; ++i) {
                const uint8_t ec = active[si].empty() ? 0 : active[si][i - n_inf];
                if (ec) {
                    edge_is_bnd[put] = (ec == 1);
                    neig[put++] = neig_of(si, i);
                }
            }
            uint32_t t = 0;
            for (uint32_t k = elemrow[si]; k < elemrow[si + 1]; ++k) {
                uint32_t b = neig_of(si, tri[si][t].a), c = neig_of(si, tri[si][t].b);
                while (b < si || c < si) {
                    ++t;
                    b = neig_of(si, tri[si][t].a);
                    c = neig_of(si, tri[si][t].b);
                }
                elem[3ull * k] = si;
                elem[3ull * k + 1] = b;
                elem[3ull * k + 2] = c;
                ++t;
            }
        }
    }
};

// --------------------------------------------------------------------------
// P1 pressure Laplacian (no reference counterpart).  Element loop over the
// alpha-shape triangles; K_ij = sum over the (at most two) triangles sharing
// edge ij of -cot(theta_k)/2, cot evaluated from the triangle's vertices in
// ascending global-id order so both endpoints see the same bits.  Diagonal =
// -(sum of the row's off-diagonals in ascending column order).  Dirichlet
// rows (node_is_bnd) become identity rows and their columns are zeroed.
// Load: b_i = f_i * sum_T |T|/3 on free rows, 0 on Dirichlet rows.
// --------------------------------------------------------------------------
template <class S> struct P1Tri {
    // cot of the angle at vertex k of (p0,p1,p2) with p sorted by global id
    static S cot_at(const S* pk, const S* pi, const S* pj) {
        const S ex = pi[0] - pk[0], ey = pi[1] - pk[1];
        const S fx = pj[0] - pk[0], fy = pj[1] - pk[1];
        const S dot = ex * fx + ey * fy;
        const S crs = std::fabs(ex * fy - ey * fx);
        return dot / crs;
    }
};

struct CsrD {
    uint32_t n = 0;
    std::vector<uint32_t> row, col;
    std::vector<double> val;
};

// nodes: Morton-ordered xy (2n), elem: 3 per triangle, bnd: node_is_bnd,
// f: nodal source.  Returns CSR with sorted columns and the load vector.
inline void assemble_p1(uint32_t n, const double* xy, uint32_t ne, const uint32_t* elem, const uint8_t* bnd,
                        const double* f, CsrD& A, std::vector<double>& b) {
    std::vector<std::vector<std::pair<uint32_t, double>>> rows(n);
    std::vector<double> mass(n, 0.0);
    for (uint32_t t = 0; t < ne; ++t) {
        uint32_t v[3] = {elem[3 * t], elem[3 * t + 1], elem[3 * t + 2]};
        std::sort(v, v + 3);
        const double* p[3] = {xy + 2 * v[0], xy + 2 * v[1], xy + 2 * v[2]};
        // cot at each vertex, canonical operand order (other two ascending)
        const double c0 = P1Tri<double>::cot_at(p[0], p[1], p[2]);
        const double c1 = P1Tri<double>::cot_at(p[1], p[0], p[2]);
        const double c2 = P1Tri<double>::cot_at(p[2], p[0], p[1]);
        const double ux = p[1][0] - p[0][0], uy = p[1][1] - p[0][1];
        const double wx = p[2][0] - p[0][0], wy = p[2][1] - p[0][1];
        const double area = 0.5 * std::fabs(ux * wy - uy * wx);
        // edge (v1,v2) opposite v0, (v0,v2) opposite v1, (v0,v1) opposite v2
        const double k12 = -0.5 * c0, k02 = -0.5 * c1, k01 = -0.5 * c2;
        rows[v[0]].push_back({v[1], k01});
        rows[v[1]].push_back({v[0], k01});
        rows[v[0]].push_back({v[2], k02});
        rows[v[2]].push_back({v[0], k02});
        rows[v[1]].push_back({v[2], k12});
        rows[v[2]].push_back({v[1], k12});
        for (int k = 0; k < 3; ++k) mass[v[k]] += area / 3.0;
    }
    A.n = n;
    A.row.assign(n + 1, 0);
    A.col.clear();
    A.val.clear();
    b.assign(n, 0.0);
    for (uint32_t i = 0; i < n; ++i) {
        auto& r = rows[i];
        // merge the (at most two) contributions per column: sum of two terms
        // is commutative so the order of arrival does not matter
        std::sort(r.begin(), r.end(), [](auto& x, auto& y) { return x.first < y.first; });
        std::vector<std::pair<uint32_t, double>> m;
        for (auto& e : r) {
            if (!m.empty() && m.back().first == e.first)
                m.back().second += e.second;
            else
                m.push_back(e);
        }
        double diag = 0.0;
        for (auto& e : m) diag -= e.second;
        // insert the diagonal at its sorted position
        bool placed = false;
        for (auto& e : m) {
            if (!placed && e.first > i) {
                A.col.push_back(i);
                A.val.push_back(bnd[i] ? 1.0 : diag);
                placed = true;
            }
            A.col.push_back(e.first);
            A.val.push_back((bnd[i] || bnd[e.first]) ? 0.0 : e.second);
        }
        if (!placed) {
            A.col.push_back(i);
            A.val.push_back(bnd[i] ? 1.0 : diag);
        }
        A.row[i + 1] = (uint32_t)A.col.size();
        b[i] = bnd[i] ? 0.0 : f[i] * mass[i];
    }
}

// --------------------------------------------------------------------------
// Jacobi-preconditioned CG (src/numerics/conjugate_gradient.cpp:61-136) with
// the stopping rule replaced by ||r||/||b|| <= rtol on the recursive
// residual, followed by a true-residual check; a restart from r = b - Ax is
// made when the true residual misses the target.
// --------------------------------------------------------------------------
inline void spmv(const CsrD& A, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)A.n; ++i) {
        double s = 0.0;
        for (uint32_t k = A.row[i]; k < A.row[i + 1]; ++k) s += A.val[k] * x[A.col[k]];
        y[i] = s;
    }
}
inline double dot(uint32_t n, const double* a, const double* b) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) s += a[i] * b[i];
    return s;
}

// Returns iterations; x in/out.  true_rel receives ||b-Ax||/||b||.
inline uint32_t jacobi_pcg(const CsrD& A, const double* b, double* x, double rtol, uint32_t maxit,
                           double* true_rel) {
    const uint32_t n = A.n;
    std::vector<double> idiag(n, 0.0), r(n), z(n), d(n), q(n);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t k = A.row[i]; k < A.row[i + 1]; ++k)
            if (A.col[k] == i) {
                idiag[i] = 1.0 / A.val[k];
                break;
            }
    const double bn = std::sqrt(dot(n, b, b));
    uint32_t it = 0;
    for (int restart = 0; restart < 8; ++restart) {
        spmv(A, x, q.data());
        for (uint32_t i = 0; i < n; ++i) {
            r[i] = b[i] - q[i];
            z[i] = r[i] * idiag[i];
            d[i] = z[i];
        }
        double rho = dot(n, r.data(), z.data());
        double rr = dot(n, r.data(), r.data());
        while (std::sqrt(rr) > rtol * bn && it < maxit) {
            spmv(A, d.data(), q.data());
            const double alpha = rho / dot(n, d.data(), q.data());
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < (int64_t)n; ++i) {
                x[i] += alpha * d[i];
                r[i] -= alpha * q[i];
                z[i] = r[i] * idiag[i];
            }
            const double rho_new = dot(n, r.data(), z.data());
            rr = dot(n, r.data(), r.data());
            const double beta = rho_new / rho;
            rho = rho_new;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < (int64_t)n; ++i) d[i] = z[i] + beta * d[i];
            ++it;
        }
        spmv(A, x, q.data());
        double tr = 0.0;
        for (uint32_t i = 0; i < n; ++i) tr += (b[i] - q[i]) * (b[i] - q[i]);
        const double rel = bn > 0 ? std::sqrt(tr) / bn : std::sqrt(tr);
        if (true_rel) *true_rel = rel;
        if (rel <= rtol || it >= maxit) break;
    }
    return it;
}

}  // namespace orc
