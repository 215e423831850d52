// CPU oracle, 3D part: orient3d / insphere (filters and exact signs) and the
// 3D alpha-shape (src/mesh/DirectAlphaShape3D.cpp:121-743).
//
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).  Pinned by the reference's 3D
// workload (testcases/IMR26/README.md:94-99: 100k glibc-rand points in the
// unit cube, alpha 0.022, FP32; the README reports 526,040 tetrahedra, the
// exact answer -- Qhull Delaunay + the circumradius filter -- is 526,042 and
// is what this restatement gives, tetrahedron for tetrahedron), the
// reference's predicates compiled in oracle/_ref (including the 425 exact
// insphere calls of that workload), exact rationals, and Delaunay invariants
// checked against Qhull.
#pragma once
#include <atomic>
#include <mutex>

#include "oracle.hpp"

namespace orc {

namespace exact {

// det | x y z 1 | over rows a, b, c, d (= Shewchuk's orient3d: the 3x3
// determinant of the differences to d).  Laplace on the ones column.
template <class S> int orient3d_sign(const S* a, const S* b, const S* c, const S* d) {
    const double P[4][3] = {{a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]}, {d[0], d[1], d[2]}};
    Acc acc;
    for (int k = 0; k < 4; ++k) {
        int r[3], m = 0;
        for (int q = 0; q < 4; ++q)
            if (q != k) r[m++] = q;
        const double cof = (k % 2 == 0) ? -1.0 : 1.0;  // (-1)^(k+3)
        // det3 of (x, y, z) over rows r0, r1, r2
        static const int perm[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {0, 2, 1}, {1, 0, 2}, {2, 1, 0}};
        for (int t = 0; t < 6; ++t) {
            const double f[3] = {P[r[perm[t][0]]][0], P[r[perm[t][1]]][1], P[r[perm[t][2]]][2]};
            acc.add_product(f, 3, cof * (t < 3 ? 1.0 : -1.0));
        }
    }
    return acc.sign();
}

// det | x y z (x^2+y^2+z^2) 1 | over rows a..e (= Shewchuk's insphere:
// positive when e lies inside the sphere through a, b, c, d with
// orient3d(a, b, c, d) > 0).  Laplace on the ones column, then on the lift
// column, the lift expanded into its three squares: 360 monomials.
template <class S> int insphere_sign(const S* a, const S* b, const S* c, const S* d, const S* e) {
    const double P[5][3] = {{a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]},
                            {d[0], d[1], d[2]}, {e[0], e[1], e[2]}};
    static const int perm[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {0, 2, 1}, {1, 0, 2}, {2, 1, 0}};
    Acc acc;
    for (int k = 0; k < 5; ++k) {
        int R[4], m = 0;
        for (int q = 0; q < 5; ++q)
            if (q != k) R[m++] = q;
        const double ck = (k % 2 == 0) ? 1.0 : -1.0;  // (-1)^(k+4)
        for (int j = 0; j < 4; ++j) {
            int r[3], t3 = 0;
            for (int q = 0; q < 4; ++q)
                if (q != j) r[t3++] = R[q];
            const double cj = (j % 2 == 0) ? -1.0 : 1.0;  // (-1)^(j+3)
            for (int t = 0; t < 6; ++t) {
                const double st = t < 3 ? 1.0 : -1.0;
                for (int comp = 0; comp < 3; ++comp) {
                    const double f[5] = {P[R[j]][comp], P[R[j]][comp], P[r[perm[t][0]]][0], P[r[perm[t][1]]][1],
                                         P[r[perm[t][2]]][2]};
                    acc.add_product(f, 5, ck * cj * st);
                }
            }
        }
    }
    return acc.sign();
}
}  // namespace exact

template <class S> struct PredConst3 {
    static constexpr S eps = Consts<S>::eps;
    static constexpr S o3derrboundA = (S)((7.0 + 56.0 * (double)eps) * (double)eps);
    static constexpr S isperrboundA = (S)((16.0 + 224.0 * (double)eps) * (double)eps);
};

// orient3d filter + exact (include/geometry/predicates.hpp:661-714); the
// exact branch returns the sign only.
template <class S> S orient3d(const S* pa, const S* pb, const S* pc, const S* pd) {
    const S adx = pa[0] - pd[0], bdx = pb[0] - pd[0], cdx = pc[0] - pd[0];
    const S ady = pa[1] - pd[1], bdy = pb[1] - pd[1], cdy = pc[1] - pd[1];
    const S adz = pa[2] - pd[2], bdz = pb[2] - pd[2], cdz = pc[2] - pd[2];
    const S bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const S cdxady = cdx * ady, adxcdy = adx * cdy;
    const S adxbdy = adx * bdy, bdxady = bdx * ady;
    const S det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    const S permanent = (std::fabs(bdxcdy) + std::fabs(cdxbdy)) * std::fabs(adz) +
                        (std::fabs(cdxady) + std::fabs(adxcdy)) * std::fabs(bdz) +
                        (std::fabs(adxbdy) + std::fabs(bdxady)) * std::fabs(cdz);
    const S errbound = PredConst3<S>::o3derrboundA * permanent;
    if ((det > errbound) || (-det > errbound)) return det;
    return (S)exact::orient3d_sign(pa, pb, pc, pd);
}

// insphere filter (predicates.hpp:985-1100): returns det, writes err.
template <class S> S insphere_filter(const S* pa, const S* pb, const S* pc, const S* pd, const S* pe, S* err) {
    const S aex = pa[0] - pe[0], bex = pb[0] - pe[0], cex = pc[0] - pe[0], dex = pd[0] - pe[0];
    const S aey = pa[1] - pe[1], bey = pb[1] - pe[1], cey = pc[1] - pe[1], dey = pd[1] - pe[1];
    const S aez = pa[2] - pe[2], bez = pb[2] - pe[2], cez = pc[2] - pe[2], dez = pd[2] - pe[2];
    const S aexbey = aex * bey, bexaey = bex * aey, ab = aexbey - bexaey;
    const S bexcey = bex * cey, cexbey = cex * bey, bc = bexcey - cexbey;
    const S cexdey = cex * dey, dexcey = dex * cey, cd = cexdey - dexcey;
    const S dexaey = dex * aey, aexdey = aex * dey, da = dexaey - aexdey;
    const S aexcey = aex * cey, cexaey = cex * aey, ac = aexcey - cexaey;
    const S bexdey = bex * dey, dexbey = dex * bey, bd = bexdey - dexbey;
    const S abc = std::fma(aez, bc, std::fma(-bez, ac, cez * ab));
    const S bcd = std::fma(bez, cd, std::fma(-cez, bd, dez * bc));
    const S cda = std::fma(cez, da, std::fma(dez, ac, aez * cd));
    const S dab = std::fma(dez, ab, std::fma(aez, bd, bez * da));
    const S alift = std::fma(aex, aex, std::fma(aey, aey, aez * aez));
    const S blift = std::fma(bex, bex, std::fma(bey, bey, bez * bez));
    const S clift = std::fma(cex, cex, std::fma(cey, cey, cez * cez));
    const S dlift = std::fma(dex, dex, std::fma(dey, dey, dez * dez));
    const S det = std::fma(dlift, abc, std::fma(-clift, dab, std::fma(blift, cda, -alift * bcd)));
    const S az = std::fabs(aez), bz = std::fabs(bez), cz = std::fabs(cez), dz = std::fabs(dez);
    const S p_aexbey = std::fabs(aexbey), p_bexaey = std::fabs(bexaey), p_bexcey = std::fabs(bexcey),
            p_cexbey = std::fabs(cexbey), p_cexdey = std::fabs(cexdey), p_dexcey = std::fabs(dexcey),
            p_dexaey = std::fabs(dexaey), p_aexdey = std::fabs(aexdey), p_aexcey = std::fabs(aexcey),
            p_cexaey = std::fabs(cexaey), p_bexdey = std::fabs(bexdey), p_dexbey = std::fabs(dexbey);
    const S permanent = std::fma(
        alift,
        std::fma(bz, (p_cexdey + p_dexcey), std::fma(cz, (p_dexbey + p_bexdey), dz * (p_bexcey + p_cexbey))),
        std::fma(blift,
                 std::fma(cz, (p_dexaey + p_aexdey), std::fma(dz, (p_aexcey + p_cexaey), az * (p_cexdey + p_dexcey))),
                 std::fma(clift,
                          std::fma(dz, (p_aexbey + p_bexaey),
                                   std::fma(az, (p_bexdey + p_dexbey), bz * (p_dexaey + p_aexdey))),
                          dlift * (std::fma(az, (p_bexcey + p_cexbey),
                                            std::fma(bz, (p_cexaey + p_aexcey), cz * (p_aexbey + p_bexaey)))))));
    *err = PredConst3<S>::isperrboundA * permanent;
    return det;
}

// --------------------------------------------------------------------------
// 3D alpha-shape (src/mesh/DirectAlphaShape3D.cpp:121-743).
// --------------------------------------------------------------------------
struct Tri8 {
    uint8_t a, b, c;
};

template <class S> struct AlphaShape3D {
    static constexpr uint32_t n_inf = 4;
    static constexpr uint32_t max_cand = 128;  // cpp:369-370 (including the infinity nodes)
    int max_cavity_iter = 100;  // cpp:519 (a parameter here only to show the cap's effect)
    uint32_t n = 0;
    int morton_bits = 21;
    int contract = 0;  // fmac mode of the inexact tests
    Lbvh<S, 3> lbvh;
    std::vector<Sphere<S, 3>> coords_m;        // Morton order + 4 infinity nodes
    std::vector<uint32_t> count;               // count pass (cpp:178-231)
    std::vector<uint32_t> nineig;              // min(4 + candidates, 128)
    std::vector<std::vector<uint32_t>> cand;   // candidate ids in traversal order
    std::vector<std::vector<Tri8>> tet;        // local tets: accepted-owned first
    std::vector<std::vector<uint8_t>> active;  // per-candidate edge counts
    std::vector<uint8_t> node_is_bnd, node_nelem, node_nelem_out;
    std::vector<uint32_t> node_nfneig;
    std::vector<uint32_t> row, elemrow, neig;
    std::vector<uint8_t> edge_is_bnd;
    std::vector<uint32_t> elem;  // 4 per tetrahedron
    uint32_t n_elems = 0, n_edges = 0;
    uint64_t exact_calls = 0, over_cap = 0;
    std::atomic<uint64_t> broken_cycles{0};
    std::atomic<uint64_t> capped_walks{0};  // cavity walks stopped by the 100-step cap
    std::vector<uint8_t> node_capped;       // 1 where a walk of this node hit the cap
    std::vector<uint32_t> node_tmax;        // largest local tet count during insertion
    bool record_exact = false;              // keep the points of every exact insphere call
    std::vector<S> exact_points;            // 15 per call, in the call's (SoS-sorted) order
    std::mutex exact_mu;

    template <class Visit> void traverse(uint32_t tid, Visit visit) const {
        int stack[64];
        int ss = 0;
        stack[ss++] = 0;
        const Sphere<S, 3> si = coords_m[tid];
        const S two_delta = Consts<S>::two_delta;
        const S rs = two_delta * si.r;
        while (ss != 0) {
            const uint32_t cur = (uint32_t)stack[--ss];
            const uint32_t sepv = (uint32_t)lbvh.sep[cur];
            const uint8_t cl = lbvh.child_is_leaf[cur];
            for (uint32_t ch = 0; ch < 2; ++ch) {
                if (!(cl & (ch + 1))) {
                    const Box<S, 3>& bb = lbvh.internal_box[sepv + ch];
                    const S dmaxx = std::fmax(si.c[0] - bb.hi[0], (S)0);
                    const S dmaxy = std::fmax(si.c[1] - bb.hi[1], (S)0);
                    const S dmaxz = std::fmax(si.c[2] - bb.hi[2], (S)0);
                    const S dminx = std::fmax(bb.lo[0] - si.c[0], (S)0);
                    const S dminy = std::fmax(bb.lo[1] - si.c[1], (S)0);
                    const S dminz = std::fmax(bb.lo[2] - si.c[2], (S)0);
                    const S dd[6] = {dmaxx, dminx, dmaxy, dminy, dmaxz, dminz};
                    const bool hit = fmac::sumsq(contract, contract & fmac::kBox, dd, 6) < rs * rs;
                    stack[ss] = (int)(sepv + ch);
                    ss += hit;
                } else if (tid != sepv + ch) {
                    const Sphere<S, 3>& sj = coords_m[sepv + ch];
                    const S ma = std::fmin(si.r, sj.r);
                    const S dxyz[3] = {si.c[0] - sj.c[0], si.c[1] - sj.c[1], si.c[2] - sj.c[2]};
                    const S sq = fmac::sumsq(contract, contract & fmac::kLeaf, dxyz, 3);
                    if (sq < (two_delta * ma) * (two_delta * ma)) visit(sepv + ch);
                }
            }
        }
    }

    void init(const std::vector<Sphere<S, 3>>& nodes) {
        n = (uint32_t)nodes.size();
        build_lbvh(lbvh, nodes, morton_bits, [](const Sphere<S, 3>& s) { return centre_box(s); });
        coords_m = lbvh.obj_m;
        // infinity nodes (cpp:144-170): avg and range of the sphere bbox, r = 1
        const Box<S, 3>& g = lbvh.global;
        S avg[3], range[3];
        for (int a = 0; a < 3; ++a) {
            avg[a] = (S)0.5 * (g.hi[a] + g.lo[a]);
            range[a] = g.hi[a] - g.lo[a];
        }
        const S mr = std::fmax(range[0], std::fmax(range[1], range[2]));
        coords_m.push_back({{avg[0] - 2 * mr, avg[1] - 2 * mr, avg[2] - 2 * mr}, (S)1});
        coords_m.push_back({{avg[0] + 2 * mr, avg[1] - 2 * mr, avg[2] - 2 * mr}, (S)1});
        coords_m.push_back({{avg[0] + 2 * mr, avg[1] + 2 * mr, avg[2] - 2 * mr}, (S)1});
        coords_m.push_back({{avg[0], avg[1], avg[2] + 2 * mr}, (S)1});
        count.assign(n, 0);
#pragma omp parallel for schedule(dynamic, 1024)
        for (uint32_t tid = 0; tid < n; ++tid) {
            uint32_t c = 0;
            traverse(tid, [&](uint32_t) { ++c; });
            count[tid] = c;
        }
    }

    const S* pt(uint32_t g) const { return coords_m[g].c; }
    uint32_t neig_of(uint32_t tid, uint32_t j) const { return j >= n_inf ? cand[tid][j - n_inf] : n + j; }

    // Local Bowyer-Watson insertion on the link of the node, cavity boundary
    // by Ray et al.'s cycle walk, then the alpha filter (cpp:380-682).
    void local_delaunay(uint32_t tid, uint64_t& nexact) {
        std::vector<Tri8> T;
        const uint32_t ncand = nineig[tid];
        T.resize(2 * ncand + 8);
        T[0] = {0, 2, 1};
        T[1] = {1, 2, 3};
        T[2] = {0, 1, 3};
        T[3] = {0, 3, 2};
        active[tid].assign(cand[tid].size(), 0);
        if (ncand - n_inf < 3) {
            node_is_bnd[tid] = 1;
            node_nelem_out[tid] = 0;
            node_nelem[tid] = 0;
            node_nfneig[tid] = 0;
            tet[tid].clear();
            return;
        }
        uint8_t cycle[max_cand];
        uint32_t Tlast = 4;
        for (uint32_t pi = n_inf; pi < ncand; ++pi) {
            const uint32_t curr = neig_of(tid, pi);
            uint32_t Rsize = 0;
            uint32_t ti = 0;
            while (ti < Tlast) {
                const Tri8 tl = T[ti];
                uint32_t i1 = tid, i2 = neig_of(tid, tl.a), i3 = neig_of(tid, tl.b), i4 = neig_of(tid, tl.c), i5 = curr;
                const S *pa = pt(i1), *pb = pt(i2), *pc = pt(i3), *pd = pt(i4), *pe = pt(i5);
                S errb;
                S det = -insphere_filter(pa, pb, pc, pd, pe, &errb);
                if (!((det > errb) || (-det > errb))) {
                    ++nexact;
                    bool swap = false;
                    auto cs = [&](uint32_t& x, uint32_t& y, const S*& px, const S*& py) {
                        if (x > y) {
                            std::swap(x, y);
                            std::swap(px, py);
                            swap = !swap;
                        }
                    };
                    cs(i1, i5, pa, pe);
                    cs(i2, i4, pb, pd);
                    cs(i1, i3, pa, pc);
                    cs(i3, i5, pc, pe);
                    cs(i1, i2, pa, pb);
                    cs(i3, i4, pc, pd);
                    cs(i2, i5, pb, pe);
                    cs(i2, i3, pb, pc);
                    cs(i4, i5, pd, pe);
                    det = (S)-exact::insphere_sign(pa, pb, pc, pd, pe);
                    if (record_exact) {
                        std::lock_guard<std::mutex> lk(exact_mu);
                        for (const S* q : {pa, pb, pc, pd, pe}) exact_points.insert(exact_points.end(), q, q + 3);
                    }
                    int depth = 0;
                    while (det == (S)0) {
                        ++depth;
                        if (depth == 1)
                            det = orient3d(pb, pc, pd, pe);
                        else if (depth == 2)
                            det = -orient3d(pa, pc, pd, pe);
                        else
                            det = swap ? (S)1 : (S)-1;
                    }
                    if (swap) det = -det;
                }
                if (det > (S)0) {
                    ++Rsize;
                    --Tlast;
                    T[ti] = T[Tlast];
                    T[Tlast] = tl;
                } else {
                    ++ti;
                }
            }
            if (Rsize) {
                for (uint32_t i = 0; i < max_cand; ++i) cycle[i] = 255;
                --Rsize;
                const Tri8 tlast = T[Tlast + Rsize];
                cycle[tlast.a] = tlast.b;
                cycle[tlast.b] = tlast.c;
                cycle[tlast.c] = tlast.a;
                uint32_t cur = Tlast;
                uint32_t first = tlast.a;
                int swap_with = (int)(Tlast + Rsize) - 1;
                int niter = 0;
                while (Rsize && niter++ < max_cavity_iter) {
                    const uint8_t tc[3] = {T[cur].a, T[cur].b, T[cur].c};
                    const uint8_t ty[3] = {cycle[tc[0]], cycle[tc[1]], cycle[tc[2]]};
                    bool inb[3], opp[3];
                    for (int e = 0; e < 3; ++e) {
                        inb[e] = ty[e] != 255;
                        opp[e] = ty[(e + 1) % 3] == tc[e];
                    }
                    bool simple = opp[0] || opp[1] || opp[2];
                    for (int e = 0; e < 3; ++e) simple &= (opp[e] || opp[(e + 1) % 3] || !inb[(e + 1) % 3]);
                    if (!simple) {
                        ++cur;
                        if ((int)cur > swap_with) cur = Tlast;
                        continue;
                    }
                    for (int e = 0; e < 3; ++e)
                        if (!opp[e]) cycle[tc[e]] = tc[(e + 1) % 3];
                    for (int e = 0; e < 3; ++e) {
                        if (opp[e] && opp[(e + 1) % 3]) {
                            if (first == tc[(e + 1) % 3]) first = ty[(e + 1) % 3];
                            cycle[tc[(e + 1) % 3]] = 255;
                        }
                    }
                    T[cur] = T[swap_with];
                    T[swap_with] = {tc[0], tc[1], tc[2]};
                    --swap_with;
                    --Rsize;
                    cur = Tlast;
                }
                if (Rsize) {
                    ++capped_walks;
                    node_capped[tid] = 1;
                }
                // one new tet per boundary edge; a walk that leaves the
                // cycle (only after the iteration cap) stops at 255 steps
                uint32_t c = first;
                int guard = 0;
                do {
                    if (c >= max_cand || ++guard > 255) {
                        ++broken_cycles;
                        break;
                    }
                    if (Tlast >= T.size()) T.resize(2 * T.size());
                    T[Tlast] = {(uint8_t)pi, (uint8_t)c, cycle[c]};
                    ++Tlast;
                    if (Tlast > node_tmax[tid]) node_tmax[tid] = Tlast;
                    c = cycle[c];
                } while (c != first);
            }
        }
        // filter (cpp:590-682)
        bool bnd = false;
        uint32_t nout = 0, nloc = 0;
        for (uint32_t i = 0; i < max_cand; ++i) cycle[i] = 0;
        uint32_t ti = 0;
        while (ti < Tlast) {
            const Tri8 tl = T[ti];
            if (tl.a < n_inf || tl.b < n_inf || tl.c < n_inf) {
                bnd = true;
                --Tlast;
                T[ti] = T[Tlast];
                continue;
            }
            uint32_t i1 = tid, i2 = neig_of(tid, tl.a), i3 = neig_of(tid, tl.b), i4 = neig_of(tid, tl.c);
            if (i1 > i3) std::swap(i1, i3);
            if (i2 > i4) std::swap(i2, i4);
            if (i1 > i2) std::swap(i1, i2);
            if (i3 > i4) std::swap(i3, i4);
            if (i2 > i3) std::swap(i2, i3);
            const Sphere<S, 3>& pa = coords_m[i1];
            const Sphere<S, 3>& pb = coords_m[i2];
            const Sphere<S, 3>& pc = coords_m[i3];
            const Sphere<S, 3>& pd = coords_m[i4];
            const S t[3] = {pa.c[0] - pd.c[0], pa.c[1] - pd.c[1], pa.c[2] - pd.c[2]};
            const S u[3] = {pb.c[0] - pd.c[0], pb.c[1] - pd.c[1], pb.c[2] - pd.c[2]};
            const S v[3] = {pc.c[0] - pd.c[0], pc.c[1] - pd.c[1], pc.c[2] - pd.c[2]};
            const int m = contract;
            const bool fz = m & fmac::kFilter;
            const S d2t = fmac::sumsq(m, fz, t, 3);
            const S d2u = fmac::sumsq(m, fz, u, 3);
            const S d2v = fmac::sumsq(m, fz, v, 3);
            const S vA = fmac::diff(m, fz, u[0], v[1], v[0], u[1]);
            const S vB = fmac::diff(m, fz, t[0], v[1], v[0], t[1]);
            const S vC = fmac::diff(m, fz, t[0], u[1], u[0], t[1]);
            const S vol = fmac::acc(fz, fmac::diff(m, fz, t[2], vA, u[2], vB), v[2], vC);
            auto cross = [&](const S* p, const S* q, S* o) {
                o[0] = fmac::diff(m, fz, p[1], q[2], p[2], q[1]);
                o[1] = fmac::diff(m, fz, p[2], q[0], p[0], q[2]);
                o[2] = fmac::diff(m, fz, p[0], q[1], p[1], q[0]);
            };
            S uv_[3], vt_[3], tu_[3];
            cross(u, v, uv_);
            cross(v, t, vt_);
            cross(t, u, tu_);
            S w[3];
            for (int a = 0; a < 3; ++a) w[a] = fmac::dot2(m, fz, d2t, uv_[a], d2u, vt_[a]);
            for (int a = 0; a < 3; ++a) w[a] = fmac::acc(fz, w[a], d2v, tu_[a]);
            const S wn = fmac::sumsq(m, fz, w, 3);
            const S rc = wn / ((S)4 * vol * vol);
            const S r1 = pa.r, r2 = pb.r, r3 = pc.r, r4 = pd.r;
            if (rc <= r1 * r1 && rc <= r2 * r2 && rc <= r3 * r3 && rc <= r4 * r4) {
                active[tid][tl.a - n_inf] += 1;
                active[tid][tl.b - n_inf] += 1;
                active[tid][tl.c - n_inf] += 1;
                cycle[tl.a]++;
                cycle[tl.b]++;
                cycle[tl.c]++;
                ++nloc;
                if (tid == i1) {
                    ++nout;
                    ++ti;
                } else {
                    --Tlast;
                    T[ti] = T[Tlast];
                    T[Tlast] = tl;
                }
            } else {
                bnd = true;
                --Tlast;
                T[ti] = T[Tlast];
            }
        }
        node_is_bnd[tid] = bnd;
        node_nelem_out[tid] = (uint8_t)nout;
        node_nelem[tid] = (uint8_t)nloc;
        uint32_t nf = 0;
        for (uint32_t i = 0; i < max_cand; ++i) nf += cycle[i] != 0;
        node_nfneig[tid] = nf;
        T.resize(nout);
        tet[tid] = T;
    }

    void compute() {
        cand.assign(n, {});
        nineig.assign(n, 0);
        uint64_t oc = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : oc)
        for (uint32_t tid = 0; tid < n; ++tid) {
            auto& c = cand[tid];
            c.reserve(count[tid]);
            traverse(tid, [&](uint32_t g) { c.push_back(g); });
            uint32_t nc = n_inf + (uint32_t)c.size();
            if (nc > max_cand) {
                ++oc;
                nc = max_cand;
            }
            nineig[tid] = nc;
        }
        over_cap = oc;
        tet.assign(n, {});
        active.assign(n, {});
        node_is_bnd.assign(n, 0);
        node_capped.assign(n, 0);
        node_tmax.assign(n, 4);
        node_nelem.assign(n, 0);
        node_nelem_out.assign(n, 0);
        node_nfneig.assign(n, 0);
        uint64_t nex = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : nex)
        for (uint32_t tid = 0; tid < n; ++tid) local_delaunay(tid, nex);
        exact_calls = nex;
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

    // ELL -> CSR (cpp:692-735)
    void compress() {
        neig.assign(n_edges, 0);
        edge_is_bnd.assign(n_edges, 0);
        elem.assign(4ull * n_elems, 0);
#pragma omp parallel for schedule(dynamic, 1024)
        for (uint32_t tid = 0; tid < n; ++tid) {
            uint32_t put = row[tid];
            for (uint32_t i = n_inf; i < nineig[tid]; ++i) {
                const uint8_t ec = active[tid].empty() ? 0 : active[tid][i - n_inf];
                if (ec) {
                    edge_is_bnd[put] = (ec <= 3);
                    neig[put++] = neig_of(tid, i);
                }
            }
            uint32_t t = 0;
            for (uint32_t k = elemrow[tid]; k < elemrow[tid + 1]; ++k, ++t) {
                elem[4ull * k] = tid;
                elem[4ull * k + 1] = neig_of(tid, tet[tid][t].a);
                elem[4ull * k + 2] = neig_of(tid, tet[tid][t].b);
                elem[4ull * k + 3] = neig_of(tid, tet[tid][t].c);
            }
        }
    }
};

// --------------------------------------------------------------------------
// P1 pressure Laplacian on tetrahedra (no reference counterpart).  Per tet,
// with its vertices in ascending global id p0..p3: e_k = p_k - p0,
// g1 = e2 x e3, g2 = e3 x e1, g3 = e1 x e2, g0 = -(g1 + g2 + g3),
// det = e1 . g1, K_ab = (g_a . g_b) / (6 |det|) (= V grad phi_a . grad phi_b),
// mass |det|/24 per vertex.  Entries are summed in increasing tet index (the
// element order), so both endpoints of an edge produce the same bits;
// diagonal = -(sum of the row's off-diagonals in ascending column order);
// Dirichlet rows identity, columns zeroed; b_i = f_i * mass_i.
// --------------------------------------------------------------------------
inline void tet_stiffness(const double* p0, const double* p1, const double* p2, const double* p3, double K[6],
                          double* vol) {
    const double e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const double e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    const double e3[3] = {p3[0] - p0[0], p3[1] - p0[1], p3[2] - p0[2]};
    auto cross = [](const double* a, const double* b, double* o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    double g[4][3];
    cross(e2, e3, g[1]);
    cross(e3, e1, g[2]);
    cross(e1, e2, g[3]);
    for (int c = 0; c < 3; ++c) g[0][c] = -((g[1][c] + g[2][c]) + g[3][c]);
    const double det = e1[0] * g[1][0] + e1[1] * g[1][1] + e1[2] * g[1][2];
    const double ad = std::fabs(det);
    const double den = 6.0 * ad;
    static const int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {1, 2, 3, 2, 3, 3};
    for (int k = 0; k < 6; ++k) {
        const double* a = g[pa[k]];
        const double* b = g[pb[k]];
        K[k] = (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) / den;
    }
    *vol = ad / 6.0;
}

inline void assemble_p1_tet(uint32_t n, const double* xyz, uint32_t ne, const uint32_t* elem, const uint8_t* bnd,
                            const double* f, CsrD& A, std::vector<double>& b) {
    std::vector<std::vector<std::pair<uint32_t, double>>> rows(n);
    std::vector<double> mass(n, 0.0);
    static const int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {1, 2, 3, 2, 3, 3};
    for (uint32_t t = 0; t < ne; ++t) {
        uint32_t v[4] = {elem[4 * t], elem[4 * t + 1], elem[4 * t + 2], elem[4 * t + 3]};
        std::sort(v, v + 4);
        double K[6], vol;
        tet_stiffness(xyz + 3 * v[0], xyz + 3 * v[1], xyz + 3 * v[2], xyz + 3 * v[3], K, &vol);
        for (int k = 0; k < 6; ++k) {
            rows[v[pa[k]]].push_back({v[pb[k]], K[k]});
            rows[v[pb[k]]].push_back({v[pa[k]], K[k]});
        }
        for (int k = 0; k < 4; ++k) mass[v[k]] += vol / 4.0;
    }
    A.n = n;
    A.row.assign(n + 1, 0);
    A.col.clear();
    A.val.clear();
    b.assign(n, 0.0);
    for (uint32_t i = 0; i < n; ++i) {
        auto& r = rows[i];
        std::stable_sort(r.begin(), r.end(), [](auto& x, auto& y) { return x.first < y.first; });
        std::vector<std::pair<uint32_t, double>> m;
        for (auto& e : r) {
            if (!m.empty() && m.back().first == e.first)
                m.back().second += e.second;
            else
                m.push_back(e);
        }
        double diag = 0.0;
        for (auto& e : m) diag -= e.second;
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

}  // namespace orc
