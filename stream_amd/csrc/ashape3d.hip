// 3D alpha-shape kernels and the AlphaShape3D C-ABI (see ashape3d.hpp).
#include <chrono>
#include <vector>

#include "ashape3d.hpp"
#include "predicates.hpp"

namespace stream_amd {

template <class S> struct Consts3D;
template <> struct Consts3D<float> {
    // 2*delta (DirectAlphaShape3D.cpp:196,335): 2 + 2^-21 in float
    static constexpr float two_delta = 2.0000004768371866248429007339848490744624f;
};
template <> struct Consts3D<double> {
    static constexpr double two_delta = 2.0 + 0x1p-50;
};

template <class S> struct Ash3Args {
    uint32_t n;
    const Sph4<S>* P;        // Morton-ordered points + infinity nodes
    const uint32_t* node;    // LBVH internal nodes
    const BoxD<S, 3>* ibox;  // LBVH internal boxes
    uint32_t* cand;
    uint32_t* tet;
    uint8_t* active;
    uint32_t* nineig;
    uint32_t* nfneig;
    uint8_t* node_is_bnd;
    uint8_t* nelem;
    uint8_t* nelem_out;
    uint32_t* queue;
    uint32_t* flags;
};

namespace {

constexpr uint32_t kEmpty = 255u;

__device__ __forceinline__ uint32_t pack_tet(uint32_t a, uint32_t b, uint32_t c) { return a | (b << 8) | (c << 16); }

// Infinity nodes (cpp:144-170) from the global sphere box, radius 1.
template <class S> __global__ void infinity3_kernel(const BoxD<S, 3>* gbox, Sph4<S>* P, uint32_t n) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const BoxD<S, 3> g = *gbox;
    S avg[3], range[3];
    for (int a = 0; a < 3; ++a) {
        avg[a] = (S)0.5 * (g.hi[a] + g.lo[a]);
        range[a] = g.hi[a] - g.lo[a];
    }
    const S mr = fmax(range[0], fmax(range[1], range[2]));
    Sph4<S> p;
    p.r = (S)1;
    p.c[0] = avg[0] - 2 * mr; p.c[1] = avg[1] - 2 * mr; p.c[2] = avg[2] - 2 * mr;
    P[n] = p;
    p.c[0] = avg[0] + 2 * mr; p.c[1] = avg[1] - 2 * mr; p.c[2] = avg[2] - 2 * mr;
    P[n + 1] = p;
    p.c[0] = avg[0] + 2 * mr; p.c[1] = avg[1] + 2 * mr; p.c[2] = avg[2] - 2 * mr;
    P[n + 2] = p;
    p.c[0] = avg[0]; p.c[1] = avg[1]; p.c[2] = avg[2] + 2 * mr;
    P[n + 3] = p;
}

// Candidate search (cpp:310-363): stack DFS, children in order, leaves
// tested when their parent is visited; the first 124 candidates are kept
// (the reference clamps its list to 128 entries including the 4 infinity
// nodes and warns), queue[1] counts clamped nodes.
//
// The DFS pushes the hit children of a node in order and pops the last one
// pushed.  Here that last child is kept in a register instead of going
// through the stack, and an internal child's node word is loaded with its
// box, so descending costs one round of loads; the stack (63 entries + the
// register = the 64-deep stack; the first 16 levels in LDS, laid out level
// by level so that the lanes' words are consecutive, the rest in private
// memory) holds node words, so a pop needs no node load.  The visit order, and so the candidate order, is
// that of the plain push/pop loop, and so is the capacity error.
constexpr int kFillLds = 16;  // stack levels held in LDS ([level][thread]); deeper ones in private memory
template <class S> __global__ __launch_bounds__(256) void fill3_kernel(Ash3Args<S> a) {
    __shared__ uint32_t lstack[kFillLds][256];
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= a.n) return;
    const Sph4<S> si = a.P[tid];
    const S two_delta = Consts3D<S>::two_delta;
    const S rs = two_delta * si.r;
    uint32_t stack[63 - kFillLds];
    int ss = 0;
    bool live = a.n >= 2;
    uint32_t w = live ? a.node[0] : 0u;
    uint32_t nloc = 0;
    bool over = false;
    for (uint32_t visits = 0; live && !over; ++visits) {
        if (visits >= a.n) {
            raise_flag(a.flags, kErrCapacity);
            break;
        }
        const uint32_t sep = w & 0x3fffffffu;
        const uint32_t cl = w >> 30;
        bool hit[2] = {false, false};
        uint32_t wc[2] = {0u, 0u};
#pragma unroll
        for (uint32_t ch = 0; ch < 2; ++ch) {
            const uint32_t c = sep + ch;
            if (!(cl & (ch + 1))) {
                const BoxD<S, 3> bb = a.ibox[c];
                wc[ch] = a.node[c];
                const S dmaxx = fmax(si.c[0] - bb.hi[0], (S)0);
                const S dmaxy = fmax(si.c[1] - bb.hi[1], (S)0);
                const S dmaxz = fmax(si.c[2] - bb.hi[2], (S)0);
                const S dminx = fmax(bb.lo[0] - si.c[0], (S)0);
                const S dminy = fmax(bb.lo[1] - si.c[1], (S)0);
                const S dminz = fmax(bb.lo[2] - si.c[2], (S)0);
                hit[ch] =
                    dmaxx * dmaxx + dminx * dminx + dmaxy * dmaxy + dminy * dminy + dmaxz * dmaxz + dminz * dminz <
                    rs * rs;
            } else if (tid != c && !over) {
                const Sph4<S> sj = a.P[c];
                const S ma = fmin(si.r, sj.r);
                const S dx = si.c[0] - sj.c[0], dy = si.c[1] - sj.c[1], dz = si.c[2] - sj.c[2];
                const S sq = dx * dx + dy * dy + dz * dz;
                if (sq < (two_delta * ma) * (two_delta * ma)) {
                    if (nloc < kCand3D)
                        a.cand[slab3(tid, nloc++, kCand3D)] = c;
                    else
                        over = true;
                }
            }
        }
        if (hit[0] && hit[1]) {
            if (ss < 63) {
                if (ss < kFillLds)
                    lstack[ss][threadIdx.x] = wc[0];
                else
                    stack[ss - kFillLds] = wc[0];
                ++ss;
                w = wc[1];
            } else {
                // a full stack drops the second child and visits the first
                raise_flag(a.flags, kErrCapacity);
                w = wc[0];
            }
        } else if (hit[0] || hit[1]) {
            w = hit[1] ? wc[1] : wc[0];
        } else if (ss != 0) {
            --ss;
            w = ss < kFillLds ? lstack[ss][threadIdx.x] : stack[ss - kFillLds];
        } else {
            live = false;
        }
    }
    if (over) atomicAdd(&a.queue[1], 1u);
    a.nineig[tid] = kNInf3D + nloc;
}

template <class S> __device__ __forceinline__ void load3(const Sph4<S>* P, uint32_t i, S* p) {
    const Sph4<S> s = P[i];
    p[0] = s.c[0];
    p[1] = s.c[1];
    p[2] = s.c[2];
}

// Local Delaunay of node tid's star + alpha filter (cpp:380-682).
// EXACT = false returns false at the first insphere whose filter cannot
// certify the sign; the node is then redone by the exact kernel.
template <class S, bool EXACT>
__device__ bool local_delaunay3(const Ash3Args<S>& a, uint32_t tid, uint8_t* cyc, double* ws) {
    const uint32_t n = a.n;
    const Sph4<S>* P = a.P;
    const uint32_t ncand = a.nineig[tid];
    const uint32_t lane = threadIdx.x & 63u;
    auto neig = [&](uint32_t j) -> uint32_t { return j < kNInf3D ? n + j : a.cand[slab3(tid, j - kNInf3D, kCand3D)]; };
    auto T = [&](uint32_t ti) -> uint32_t& { return a.tet[slab3(tid, ti, kTet3D)]; };
    auto C = [&](uint32_t v) -> uint8_t& { return cyc[v * 64u + lane]; };

    T(0) = pack_tet(0, 2, 1);
    T(1) = pack_tet(1, 2, 3);
    T(2) = pack_tet(0, 1, 3);
    T(3) = pack_tet(0, 3, 2);
    if (ncand - kNInf3D < 3) {
        for (uint32_t j = 0; j + kNInf3D < ncand; ++j) a.active[slab3(tid, j, kCand3D)] = 0;
        a.node_is_bnd[tid] = 1;
        a.nelem_out[tid] = 0;
        a.nelem[tid] = 0;
        a.nfneig[tid] = 0;
        return true;
    }
    S pa0[3];
    load3(P, tid, pa0);
    uint32_t Tlast = 4;
    bool over = false;
    for (uint32_t pi = kNInf3D; pi < ncand; ++pi) {
        const uint32_t curr = neig(pi);
        S pe0[3];
        load3(P, curr, pe0);
        uint32_t Rsize = 0;
        if constexpr (!EXACT) {
            // Fast path (filters only; a node whose filter cannot decide is
            // redone by the exact kernel, which takes the sequential loop below).
            // Conflict flags of the current tetrahedra with the new point, bit
            // k of word k >> 6 (f0..f3) for position k: every tetrahedron is
            // tested once, two per step so that their loads overlap.
            uint64_t f0 = 0, f1 = 0, f2 = 0, f3 = 0;
            auto flag = [&](uint32_t k) -> bool {
                const uint64_t w = k < 64 ? f0 : (k < 128 ? f1 : (k < 192 ? f2 : f3));
                return (w >> (k & 63u)) & 1u;
            };
            auto set_flag = [&](uint32_t k, bool v) {
                const uint64_t m = 1ull << (k & 63u);
                const uint32_t q = k >> 6;
                f0 = q == 0 ? (v ? f0 | m : f0 & ~m) : f0;
                f1 = q == 1 ? (v ? f1 | m : f1 & ~m) : f1;
                f2 = q == 2 ? (v ? f2 | m : f2 & ~m) : f2;
                f3 = q == 3 ? (v ? f3 | m : f3 & ~m) : f3;
            };
            // first flagged position >= k (256 when none)
            auto next_flag = [&](uint32_t k) -> uint32_t {
#pragma unroll
                for (uint32_t q = 0; q < 4; ++q) {
                    if (k < q * 64 + 64) {
                        uint64_t m = q == 0 ? f0 : (q == 1 ? f1 : (q == 2 ? f2 : f3));
                        if (k > q * 64) m &= ~0ull << (k - q * 64);
                        if (m) return q * 64 + (uint32_t)__builtin_ctzll(m);
                    }
                }
                return 256u;
            };
            // point ids of the pair at t0 (the second repeats the first past
            // the end); the next pair's ids load while the current pair is tested
            auto pair_ids = [&](uint32_t t0, uint32_t (&o)[2][3]) {
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const uint32_t tl = T(u == 0 || t0 + 1 >= Tlast ? t0 : t0 + 1);
                    o[u][0] = neig(tl & 255u);
                    o[u][1] = neig((tl >> 8) & 255u);
                    o[u][2] = neig((tl >> 16) & 255u);
                }
            };
            uint32_t ids[2][3];
            if (Tlast) pair_ids(0, ids);
            for (uint32_t t0 = 0; t0 < Tlast; t0 += 2) {
                const bool two = t0 + 1 < Tlast;
                S pb[2][3], pc[2][3], pd[2][3];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    load3(P, ids[u][0], pb[u]);
                    load3(P, ids[u][1], pc[u]);
                    load3(P, ids[u][2], pd[u]);
                }
                uint32_t nids[2][3];
                const bool more = t0 + 2 < Tlast;
                if (more) pair_ids(t0 + 2, nids);
                S det[2], err[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) det[u] = -insphere_filter(pa0, pb[u], pc[u], pd[u], pe0, &err[u]);
                bool pos[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    pos[u] = det[u] > (S)0;
                    if (u == 1 && !two) break;
                    if ((det[u] > err[u]) || (-det[u] > err[u])) continue;
                    if constexpr (sizeof(S) == sizeof(double)) {
                        return false;
                    } else {
                        // FP32 inputs: the FP64 filter on the same (exactly
                        // widened) coordinates certifies most of the signs the
                        // FP32 filter cannot; the rest go to the exact kernel
                        double qa[3], qb[3], qc[3], qd[3], qe[3];
#pragma unroll
                        for (int v = 0; v < 3; ++v) {
                            qa[v] = pa0[v];
                            qb[v] = pb[u][v];
                            qc[v] = pc[u][v];
                            qd[v] = pd[u][v];
                            qe[v] = pe0[v];
                        }
                        double e2;
                        const double d2 = -insphere_filter(qa, qb, qc, qd, qe, &e2);
                        if (!((d2 > e2) || (-d2 > e2))) return false;
                        pos[u] = d2 > 0.0;
                    }
                }
                if (pos[0]) set_flag(t0, true);
                if (two && pos[1]) set_flag(t0 + 1, true);
                if (more) {
#pragma unroll
                    for (int u = 0; u < 2; ++u)
#pragma unroll
                        for (int v = 0; v < 3; ++v) ids[u][v] = nids[u][v];
                }
            }
            // The cavity: the sequential order of cpp:421-495 (a conflicting
            // tetrahedron is swapped with the last one, which is tested next at
            // the same position) replayed from the flags.
            for (uint32_t ti = next_flag(0); ti < Tlast; ti = next_flag(ti)) {
                ++Rsize;
                --Tlast;
                const uint32_t tl = T(ti);
                T(ti) = T(Tlast);
                T(Tlast) = tl;
                set_flag(ti, flag(Tlast));
                set_flag(Tlast, true);
            }
        } else {
            uint32_t ti = 0;
            while (ti < Tlast) {
                const uint32_t tl = T(ti);
                const uint32_t la = tl & 255u, lb = (tl >> 8) & 255u, lc = (tl >> 16) & 255u;
                uint32_t i2 = neig(la), i3 = neig(lb), i4 = neig(lc);
                S pb[3], pc[3], pd[3];
                load3(P, i2, pb);
                load3(P, i3, pc);
                load3(P, i4, pd);
                S err;
                S det = -insphere_filter(pa0, pb, pc, pd, pe0, &err);
                if (!((det > err) || (-det > err))) {
                    if constexpr (!EXACT) {
                        return false;
                    } else {
                        uint32_t i1 = tid, i5 = curr;
                        S qa[3] = {pa0[0], pa0[1], pa0[2]}, qe[3] = {pe0[0], pe0[1], pe0[2]};
                        S* p1 = qa;
                        S* p2 = pb;
                        S* p3 = pc;
                        S* p4 = pd;
                        S* p5 = qe;
                        bool swap = false;
                        auto cs = [&](uint32_t& x, uint32_t& y, S*& px, S*& py) {
                            if (x > y) {
                                const uint32_t t = x;
                                x = y;
                                y = t;
                                S* u = px;
                                px = py;
                                py = u;
                                swap = !swap;
                            }
                        };
                        cs(i1, i5, p1, p5);
                        cs(i2, i4, p2, p4);
                        cs(i1, i3, p1, p3);
                        cs(i3, i5, p3, p5);
                        cs(i1, i2, p1, p2);
                        cs(i3, i4, p3, p4);
                        cs(i2, i5, p2, p5);
                        cs(i2, i3, p2, p3);
                        cs(i4, i5, p4, p5);
                        det = (S)-insphere_exact(p1, p2, p3, p4, p5, ws, (int)kExactWs3D, &over);
                        int depth = 0;
                        while (det == (S)0) {
                            ++depth;
                            if (depth == 1) {
                                bool c1;
                                det = orient3d_filter(p2, p3, p4, p5, &c1);
                                if (!c1) det = (S)orient3d_exact(p2, p3, p4, p5, ws, (int)kExactWs3D, &over);
                            } else if (depth == 2) {
                                bool c2;
                                det = orient3d_filter(p1, p3, p4, p5, &c2);
                                if (!c2) det = (S)orient3d_exact(p1, p3, p4, p5, ws, (int)kExactWs3D, &over);
                                det = -det;
                            } else {
                                det = swap ? (S)1 : (S)-1;
                            }
                        }
                        if (swap) det = -det;
                    }
                }
                if (det > (S)0) {
                    ++Rsize;
                    --Tlast;
                    T(ti) = T(Tlast);
                    T(Tlast) = tl;
                } else {
                    ++ti;
                }
            }
        }
        if (Rsize) {
            // the star's tetrahedra use local ids below pi only
            for (uint32_t v = 0; v < pi; ++v) C(v) = (uint8_t)kEmpty;
            --Rsize;
            const uint32_t tlst = T(Tlast + Rsize);
            const uint32_t ta = tlst & 255u, tb = (tlst >> 8) & 255u, tc3 = (tlst >> 16) & 255u;
            C(ta) = (uint8_t)tb;
            C(tb) = (uint8_t)tc3;
            C(tc3) = (uint8_t)ta;
            uint32_t cur = Tlast;
            uint32_t first = ta;
            int swap_with = (int)(Tlast + Rsize) - 1;
            int niter = 0;
            while (Rsize && niter++ < 100) {
                const uint32_t tw = T(cur);
                const uint32_t tcv[3] = {tw & 255u, (tw >> 8) & 255u, (tw >> 16) & 255u};
                const uint32_t ty[3] = {C(tcv[0]), C(tcv[1]), C(tcv[2])};
                bool inb[3], opp[3];
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    inb[e] = ty[e] != kEmpty;
                    opp[e] = ty[(e + 1) % 3] == tcv[e];
                }
                bool simple = opp[0] || opp[1] || opp[2];
#pragma unroll
                for (int e = 0; e < 3; ++e) simple &= (opp[e] || opp[(e + 1) % 3] || !inb[(e + 1) % 3]);
                if (!simple) {
                    ++cur;
                    if ((int)cur > swap_with) cur = Tlast;
                    continue;
                }
#pragma unroll
                for (int e = 0; e < 3; ++e)
                    if (!opp[e]) C(tcv[e]) = (uint8_t)tcv[(e + 1) % 3];
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    if (opp[e] && opp[(e + 1) % 3]) {
                        if (first == tcv[(e + 1) % 3]) first = ty[(e + 1) % 3];
                        C(tcv[(e + 1) % 3]) = (uint8_t)kEmpty;
                    }
                }
                T(cur) = T((uint32_t)swap_with);
                T((uint32_t)swap_with) = tw;
                --swap_with;
                --Rsize;
                cur = Tlast;
            }
            uint32_t c = first;
            int guard = 0;
            do {
                if (c >= kCycle3D || ++guard > 255 || Tlast >= kTet3D) {
                    raise_flag(a.flags, kErrCapacity);
                    break;
                }
                const uint32_t nx = C(c);
                T(Tlast++) = pack_tet(pi, c, nx);
                c = nx;
            } while (c != first);
        }
    }
    if (over) raise_flag(a.flags, kErrCapacity);
    // alpha filter (cpp:590-682)
    bool bnd = false;
    uint32_t nout = 0, nloc = 0;
    for (uint32_t v = 0; v < ncand; ++v) C(v) = 0;
    uint32_t ti = 0;
    while (ti < Tlast) {
        const uint32_t tl = T(ti);
        const uint32_t la = tl & 255u, lb = (tl >> 8) & 255u, lc = (tl >> 16) & 255u;
        if (la < kNInf3D || lb < kNInf3D || lc < kNInf3D) {
            bnd = true;
            --Tlast;
            T(ti) = T(Tlast);
            continue;
        }
        uint32_t i1 = tid, i2 = neig(la), i3 = neig(lb), i4 = neig(lc), t;
        if (i1 > i3) { t = i3; i3 = i1; i1 = t; }
        if (i2 > i4) { t = i4; i4 = i2; i2 = t; }
        if (i1 > i2) { t = i2; i2 = i1; i1 = t; }
        if (i3 > i4) { t = i4; i4 = i3; i3 = t; }
        if (i2 > i3) { t = i3; i3 = i2; i2 = t; }
        const Sph4<S> qa = P[i1], qb = P[i2], qc = P[i3], qd = P[i4];
        const S tv[3] = {qa.c[0] - qd.c[0], qa.c[1] - qd.c[1], qa.c[2] - qd.c[2]};
        const S uv[3] = {qb.c[0] - qd.c[0], qb.c[1] - qd.c[1], qb.c[2] - qd.c[2]};
        const S vv[3] = {qc.c[0] - qd.c[0], qc.c[1] - qd.c[1], qc.c[2] - qd.c[2]};
        const S d2t = tv[0] * tv[0] + tv[1] * tv[1] + tv[2] * tv[2];
        const S d2u = uv[0] * uv[0] + uv[1] * uv[1] + uv[2] * uv[2];
        const S d2v = vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2];
        const S vol = tv[2] * (uv[0] * vv[1] - vv[0] * uv[1]) - uv[2] * (tv[0] * vv[1] - vv[0] * tv[1]) +
                      vv[2] * (tv[0] * uv[1] - uv[0] * tv[1]);
        S w[3];
        {
            const S x0 = uv[1] * vv[2] - uv[2] * vv[1], x1 = uv[2] * vv[0] - uv[0] * vv[2],
                    x2 = uv[0] * vv[1] - uv[1] * vv[0];
            w[0] = d2t * x0;
            w[1] = d2t * x1;
            w[2] = d2t * x2;
        }
        {
            const S x0 = vv[1] * tv[2] - vv[2] * tv[1], x1 = vv[2] * tv[0] - vv[0] * tv[2],
                    x2 = vv[0] * tv[1] - vv[1] * tv[0];
            w[0] += d2u * x0;
            w[1] += d2u * x1;
            w[2] += d2u * x2;
        }
        {
            const S x0 = tv[1] * uv[2] - tv[2] * uv[1], x1 = tv[2] * uv[0] - tv[0] * uv[2],
                    x2 = tv[0] * uv[1] - tv[1] * uv[0];
            w[0] += d2v * x0;
            w[1] += d2v * x1;
            w[2] += d2v * x2;
        }
        const S wn = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
        const S rc = wn / ((S)4 * vol * vol);
        if (rc <= qa.r * qa.r && rc <= qb.r * qb.r && rc <= qc.r * qc.r && rc <= qd.r * qd.r) {
            C(la) = (uint8_t)(C(la) + 1);
            C(lb) = (uint8_t)(C(lb) + 1);
            C(lc) = (uint8_t)(C(lc) + 1);
            ++nloc;
            if (tid == i1) {
                ++nout;
                ++ti;
            } else {
                --Tlast;
                T(ti) = T(Tlast);
                T(Tlast) = tl;
            }
        } else {
            bnd = true;
            --Tlast;
            T(ti) = T(Tlast);
        }
    }
    // per candidate, the number of kept tetrahedra using it (counted in
    // LDS, stored once; the infinity nodes are in no kept tetrahedron)
    uint32_t nf = 0;
    for (uint32_t v = kNInf3D; v < ncand; ++v) {
        const uint8_t c = C(v);
        a.active[slab3(tid, v - kNInf3D, kCand3D)] = c;
        nf += c != 0;
    }
    a.node_is_bnd[tid] = bnd;
    a.nelem_out[tid] = (uint8_t)nout;
    a.nelem[tid] = (uint8_t)nloc;
    a.nfneig[tid] = nf;
    return true;
}

// Four wavefronts per block (one LDS cycle buffer each), one node per lane.
// Lanes keep consecutive (Morton-ordered) nodes: neighbouring nodes share
// most candidates, so their coordinate loads hit the same cache lines.
constexpr uint32_t kDelaunay3Block = 256;
template <class S> __global__ __launch_bounds__(kDelaunay3Block) void delaunay3_kernel(Ash3Args<S> a) {
    __shared__ uint8_t cyc[kDelaunay3Block / 64][kCycle3D * 64];
    const uint32_t tid = blockIdx.x * kDelaunay3Block + threadIdx.x;
    if (tid >= a.n) return;
    if (!local_delaunay3<S, false>(a, tid, cyc[threadIdx.x >> 6], nullptr)) {
        const uint32_t k = atomicAdd(&a.queue[0], 1u);
        a.queue[2 + k] = tid;
    }
}

// One queued node per wavefront (lane 0; the block is one wavefront): the
// exact evaluations of different nodes happen at different points of their
// insertions, so nodes sharing a wavefront would run their exact stages one
// after another.  The expansion workspace lives in LDS.
template <class S> __global__ __launch_bounds__(64) void delaunay3_exact_kernel(Ash3Args<S> a) {
    __shared__ uint8_t cyc[kCycle3D * 64];
    __shared__ double ws[kExactWs3D];
    if (threadIdx.x != 0) return;
    const uint32_t cnt = a.queue[0];
    for (uint32_t k = blockIdx.x; k < cnt; k += gridDim.x) local_delaunay3<S, true>(a, a.queue[2 + k], cyc, ws);
}

// Slab -> CSR (cpp:692-735): edges with count <= 3 are boundary edges; the
// first nelem_out local tetrahedra are the ones this node owns.
template <class S>
__global__ __launch_bounds__(256) void compress3_kernel(Ash3Args<S> a, const uint32_t* __restrict__ row,
                                                        const uint32_t* __restrict__ elemrow,
                                                        uint32_t* __restrict__ neig, uint8_t* __restrict__ edge_is_bnd,
                                                        uint32_t* __restrict__ elem) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= a.n) return;
    const uint32_t n = a.n;
    auto gneig = [&](uint32_t j) -> uint32_t { return j < kNInf3D ? n + j : a.cand[slab3(tid, j - kNInf3D, kCand3D)]; };
    uint32_t put = row[tid];
    const uint32_t nn = a.nineig[tid];
    for (uint32_t i = kNInf3D; i < nn; ++i) {
        const uint8_t ec = a.active[slab3(tid, i - kNInf3D, kCand3D)];
        if (ec) {
            edge_is_bnd[put] = (ec <= 3);
            neig[put] = gneig(i);
            ++put;
        }
    }
    uint32_t t = 0;
    for (uint32_t k = elemrow[tid]; k < elemrow[tid + 1]; ++k, ++t) {
        const uint32_t tl = a.tet[slab3(tid, t, kTet3D)];
        elem[4ull * k] = tid;
        elem[4ull * k + 1] = gneig(tl & 255u);
        elem[4ull * k + 2] = gneig((tl >> 8) & 255u);
        elem[4ull * k + 3] = gneig((tl >> 16) & 255u);
    }
}

template <class S> Ash3Args<S> make_args3(Ash3D<S>& A) {
    Ash3Args<S> a;
    a.n = A.n;
    a.P = A.lbvh.obj_m.p;
    a.node = A.lbvh.node.p;
    a.ibox = A.lbvh.ibox.p;
    a.cand = A.cand.p;
    a.tet = A.tet.p;
    a.active = A.active.p;
    a.nineig = A.nineig.p;
    a.nfneig = A.nfneig.p;
    a.node_is_bnd = A.node_is_bnd.p;
    a.nelem = A.nelem.p;
    a.nelem_out = A.nelem_out.p;
    a.queue = A.queue.p;
    a.flags = ctx().d_flags;
    return a;
}

}  // namespace

template <class S> void Ash3D<S>::set_nodes_host(uint32_t n_, const S* xyz, const S* alpha) {
    std::vector<Sph4<S>> h(n_);
    for (uint32_t i = 0; i < n_; ++i) {
        h[i].c[0] = xyz[3ull * i];
        h[i].c[1] = xyz[3ull * i + 1];
        h[i].c[2] = xyz[3ull * i + 2];
        h[i].r = alpha[i];
    }
    n = n_;
    coords.resize(n ? n : 1);
    h2d(coords.p, h.data(), n);
    inited = computed = false;
}

template <class S> void Ash3D<S>::set_nodes_device(uint32_t n_, const Sph4<S>* d) {
    n = n_;
    coords.resize(n ? n : 1);
    d2d(coords.p, d, n);
    inited = computed = false;
}

template <class S> void Ash3D<S>::init(hipStream_t s) {
    if (n < 4) throw std::invalid_argument("AlphaShape3D needs at least 4 nodes");
    lbvh.build(coords.p, n, bits, kLeafCentre, kNInf3D, s);
    hipLaunchKernelGGL((infinity3_kernel<S>), dim3(1), dim3(64), 0, s, lbvh.gbox.p, lbvh.obj_m.p, n);
    cand.resize(slab3_words(n, kCand3D));
    tet.resize(slab3_words(n, kTet3D));
    active.resize(slab3_words(n, kCand3D));
    nineig.resize(n);
    nfneig.resize(n);
    node_is_bnd.resize(n);
    nelem.resize(n);
    nelem_out.resize(n);
    row.resize(n + 1);
    elemrow.resize(n + 1);
    queue.resize((size_t)n + 2);
    scan_ws.resize(scan_workspace_words(n));
    inited = true;
    computed = false;
}

template <class S> void Ash3D<S>::compute(hipStream_t s) {
    if (!inited) throw std::invalid_argument("AlphaShape3D_compute called before init");
    Ash3Args<S> a = make_args3(*this);
    section_mark(s, kMarkCompute0);
    fill<uint32_t>(queue.p, 2, 0u, s);
    hipLaunchKernelGGL((fill3_kernel<S>), dim3(grid_for(n, 256)), dim3(256), 0, s, a);
    section_mark(s, kMarkFill);
    hipLaunchKernelGGL((delaunay3_kernel<S>), dim3((n + kDelaunay3Block - 1) / kDelaunay3Block),
                       dim3(kDelaunay3Block), 0, s, a);
    section_mark(s, kMarkDelaunayMain);
    hipLaunchKernelGGL((delaunay3_exact_kernel<S>), dim3(kExactWaves3D), dim3(64), 0, s, a);
    section_mark(s, kMarkDelaunay);
    exclusive_scan_u8(nelem_out.p, elemrow.p, n, scan_ws.p, s);
    exclusive_scan_u32(nfneig.p, row.p, n, scan_ws.p, s);
    section_mark(s, kMarkScans);
    n_elems_ = read_scalar(elemrow.p + n);
    n_edges_ = read_scalar(row.p + n);
    neig.resize(n_edges_ ? n_edges_ : 1);
    edge_is_bnd.resize(n_edges_ ? n_edges_ : 1);
    elemglob.resize(4ull * (n_elems_ ? n_elems_ : 1));
    a = make_args3(*this);
    section_mark(s, kMarkCompress0);
    hipLaunchKernelGGL((compress3_kernel<S>), dim3(grid_for(n, 256)), dim3(256), 0, s, a, row.p, elemrow.p, neig.p,
                       edge_is_bnd.p, elemglob.p);
    section_mark(s, kMarkCompress);
    computed = true;
}

template <class S> uint32_t Ash3D<S>::n_exact_nodes() { return computed ? read_scalar(queue.p) : 0u; }
template <class S> uint32_t Ash3D<S>::n_over_cap() { return computed ? read_scalar(queue.p + 1) : 0u; }

template struct Ash3D<float>;
template struct Ash3D<double>;

}  // namespace stream_amd

// ---------------------------------------------------------------------------
// C-ABI: the reference's AlphaShape3D_* surface (DirectAlphaShape3D.cpp:
// 743-814) plus FP64 and query extensions, as for AlphaShape2D.
// ---------------------------------------------------------------------------
using namespace stream_amd;

struct AlphaShape3D {
    int bits = 21;
    bool f64 = false;
    Ash3D<float> a32;
    Ash3D<double> a64;
    int32_t status = 0;
};

namespace {
template <class F> void run_void3(AlphaShape3D* h, F&& f) {
    if (!h) {
        set_last_error("null AlphaShape3D handle");
        return;
    }
    h->status = guarded(f);
}
template <class F> int32_t run_status3(AlphaShape3D* h, F&& f) {
    if (!h) {
        set_last_error("null AlphaShape3D handle");
        return kErrArgument;
    }
    h->status = guarded(f);
    return h->status;
}
void check_flags3() {
    const uint32_t f = read_scalar(ctx().d_flags + kFlagError);
    if (f) {
        fill<uint32_t>(ctx().d_flags, 1, 0u, ctx().stream);
        throw std::runtime_error(f == kErrCapacity ? "device capacity overflow (3D alpha-shape)"
                                                   : "device error flag raised");
    }
}
}  // namespace

extern "C" {

AlphaShape3D* AlphaShape3D_create(void) {
    try {
        (void)ctx();
        return new AlphaShape3D();
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return nullptr;
    }
}

void AlphaShape3D_destroy(AlphaShape3D* h) { delete h; }

// Morton resolution (bits per axis, 1..21) of the next init.
int32_t AlphaShape3D_set_morton_bits(AlphaShape3D* h, int32_t bits) {
    return run_status3(h, [&] {
        if (bits < 1 || bits > 21) throw std::invalid_argument("3D Morton bits per axis must be in [1, 21]");
        h->bits = bits;
    });
}

// FP32 nodes: coords row-major [3N], alpha [N] (reference signature).
void AlphaShape3D_set_nodes(AlphaShape3D* h, uint32_t nnodes, const float* coords, const float* alpha) {
    run_void3(h, [&] {
        h->f64 = false;
        h->a32.set_nodes_host(nnodes, coords, alpha);
    });
}

int32_t AlphaShape3D_set_nodes_f64(AlphaShape3D* h, uint32_t nnodes, const double* coords, const double* alpha) {
    return run_status3(h, [&] {
        h->f64 = true;
        h->a64.set_nodes_host(nnodes, coords, alpha);
    });
}

// FP64 nodes already in device memory, rows (x, y, z, alpha).
int32_t AlphaShape3D_set_nodes_device_f64(AlphaShape3D* h, uint32_t nnodes, const double* d_xyzr) {
    return run_status3(h, [&] {
        h->f64 = true;
        h->a64.set_nodes_device(nnodes, reinterpret_cast<const Sph4<double>*>(d_xyzr));
    });
}

void AlphaShape3D_init(AlphaShape3D* h) {
    run_void3(h, [&] {
        if (h->f64) {
            h->a64.bits = h->bits;
            h->a64.init(ctx().stream);
        } else {
            h->a32.bits = h->bits;
            h->a32.init(ctx().stream);
        }
    });
}

void AlphaShape3D_compute(AlphaShape3D* h) {
    run_void3(h, [&] {
        if (h->f64)
            h->a64.compute(ctx().stream);
        else
            h->a32.compute(ctx().stream);
        check_flags3();
    });
}

uint32_t AlphaShape3D_get_nelem(const AlphaShape3D* hc) {
    AlphaShape3D* h = const_cast<AlphaShape3D*>(hc);
    uint32_t r = 0;
    run_void3(h, [&] { r = h->f64 ? h->a64.n_elems() : h->a32.n_elems(); });
    return r;
}

// Tetrahedra [4*nelem], Morton ids, each output by its lowest id.
void AlphaShape3D_get_elem(const AlphaShape3D* hc, uint32_t* elems) {
    AlphaShape3D* h = const_cast<AlphaShape3D*>(hc);
    run_void3(h, [&] {
        const uint32_t ne = h->f64 ? h->a64.n_elems() : h->a32.n_elems();
        d2h(elems, h->f64 ? h->a64.elemglob.p : h->a32.elemglob.p, 4ull * ne);
    });
}

void AlphaShape3D_get_ordered_nodes(AlphaShape3D* h, float* nodes) {
    run_void3(h, [&] {
        if (h->f64) throw std::invalid_argument("handle holds FP64 nodes: use AlphaShape3D_get_ordered_nodes_f64");
        std::vector<Sph4<float>> v(h->a32.n);
        d2h(v.data(), h->a32.lbvh.obj_m.p, h->a32.n);
        for (uint32_t i = 0; i < h->a32.n; ++i)
            for (int c = 0; c < 3; ++c) nodes[3 * i + c] = v[i].c[c];
    });
}

int32_t AlphaShape3D_get_ordered_nodes_f64(AlphaShape3D* h, double* nodes) {
    return run_status3(h, [&] {
        if (!h->f64) throw std::invalid_argument("handle holds FP32 nodes");
        std::vector<Sph4<double>> v(h->a64.n);
        d2h(v.data(), h->a64.lbvh.obj_m.p, h->a64.n);
        for (uint32_t i = 0; i < h->a64.n; ++i)
            for (int c = 0; c < 3; ++c) nodes[3 * i + c] = v[i].c[c];
    });
}

uint32_t AlphaShape3D_get_nedge(const AlphaShape3D* hc) {
    AlphaShape3D* h = const_cast<AlphaShape3D*>(hc);
    uint32_t r = 0;
    run_void3(h, [&] { r = h->f64 ? h->a64.n_edges() : h->a32.n_edges(); });
    return r;
}

// Edge CSR (AlphaShape3D::getEdge, cpp:90-98), edge flags (count <= 3) and
// node flags (getBoundaryNodes, cpp:101-105); any pointer may be null.
int32_t AlphaShape3D_get_edges(const AlphaShape3D* hc, uint32_t* row, uint32_t* neig, uint8_t* edge_is_bnd,
                               uint8_t* node_is_bnd) {
    AlphaShape3D* h = const_cast<AlphaShape3D*>(hc);
    return run_status3(h, [&] {
        auto go = [&](auto& A) {
            const uint32_t ne = A.n_edges();
            if (row) d2h(row, A.row.p, A.n + 1);
            if (neig) d2h(neig, A.neig.p, ne);
            if (edge_is_bnd) d2h(edge_is_bnd, A.edge_is_bnd.p, ne);
            if (node_is_bnd) d2h(node_is_bnd, A.node_is_bnd.p, A.n);
        };
        if (h->f64)
            go(h->a64);
        else
            go(h->a32);
    });
}

// perm[i] = input index of Morton slot i (getPermutation, cpp:75-80).
int32_t AlphaShape3D_get_permutation(const AlphaShape3D* hc, uint32_t* perm) {
    AlphaShape3D* h = const_cast<AlphaShape3D*>(hc);
    return run_status3(h, [&] {
        if (h->f64)
            d2h(perm, h->a64.lbvh.perm, h->a64.n);
        else
            d2h(perm, h->a32.lbvh.perm, h->a32.n);
    });
}

// Nodes that needed the exact stage; nodes whose candidate list was clamped.
uint32_t AlphaShape3D_get_nexact(const AlphaShape3D* hc) {
    AlphaShape3D* h = const_cast<AlphaShape3D*>(hc);
    uint32_t r = 0;
    run_void3(h, [&] { r = h->f64 ? h->a64.n_exact_nodes() : h->a32.n_exact_nodes(); });
    return r;
}
uint32_t AlphaShape3D_get_nclamped(const AlphaShape3D* hc) {
    AlphaShape3D* h = const_cast<AlphaShape3D*>(hc);
    uint32_t r = 0;
    run_void3(h, [&] { r = h->f64 ? h->a64.n_over_cap() : h->a32.n_over_cap(); });
    return r;
}

int32_t AlphaShape3D_last_status(const AlphaShape3D* h) { return h ? h->status : kErrArgument; }


// Nodal fields between input order and Morton order (permuteForward /
// permuteBackward, include/mesh/DirectAlphaShape3D.hpp): forward out[i] =
// in[perm[i]] gives the field in the Morton order of the outputs, backward
// out[perm[i]] = in[i] returns it to input order.  n elements of
// elem_bytes each; host arrays (copied) or device buffers (_device, on the
// library stream, not synchronised).  Needs init (the permutation).
}  // extern "C"

namespace {
template <class F> int32_t permute_call(const AlphaShape3D* hc, F&& f) {
    AlphaShape3D* h = const_cast<AlphaShape3D*>(hc);
    return run_status3(h, [&] {
        const bool ok = h->f64 ? h->a64.inited : h->a32.inited;
        if (!ok) throw std::invalid_argument("AlphaShape3D: permutation needs init (set_nodes)");
        const uint32_t* perm = h->f64 ? h->a64.lbvh.perm : h->a32.lbvh.perm;
        const uint32_t n = h->f64 ? h->a64.n : h->a32.n;
        f(perm, n);
    });
}
}  // namespace

extern "C" {

int32_t AlphaShape3D_permute_forward(const AlphaShape3D* h, const void* in, void* out, uint32_t elem_bytes) {
    return permute_call(h, [&](const uint32_t* perm, uint32_t n) {
        permute_host_elements(in, out, perm, n, elem_bytes, true);
    });
}
int32_t AlphaShape3D_permute_backward(const AlphaShape3D* h, const void* in, void* out, uint32_t elem_bytes) {
    return permute_call(h, [&](const uint32_t* perm, uint32_t n) {
        permute_host_elements(in, out, perm, n, elem_bytes, false);
    });
}
int32_t AlphaShape3D_permute_forward_device(const AlphaShape3D* h, const void* d_in, void* d_out, uint32_t elem_bytes) {
    return permute_call(h, [&](const uint32_t* perm, uint32_t n) {
        permute_elements(d_in, d_out, perm, n, elem_bytes, true, ctx().stream);
    });
}
int32_t AlphaShape3D_permute_backward_device(const AlphaShape3D* h, const void* d_in, void* d_out, uint32_t elem_bytes) {
    return permute_call(h, [&](const uint32_t* perm, uint32_t n) {
        permute_elements(d_in, d_out, perm, n, elem_bytes, false, ctx().stream);
    });
}


// One init + compute timed as testcases/IMR26/uniform3d.cpp times them:
// ms[0] tlbvh (LBVH build), ms[1] tashape (candidate search, local
// Delaunay + filter, element/edge scans), ms[2] tcompress (compaction to
// CSR), ms[3] the host wall time of both calls, ms[4] the GPU time from the
// build's start to the compaction's end; within tashape, ms[5] the
// candidate search, ms[6] the local Delaunay + filter of every node, ms[7]
// the exact redo of the nodes whose filters could not decide a sign.  GPU times come from HIP events
// on the library stream between the launches.
int32_t AlphaShape3D_timed_run(AlphaShape3D* h, double* ms) {
    return run_status3(h, [&] {
        if (!ms) throw std::invalid_argument("AlphaShape3D_timed_run: null ms");
        struct Disarm {
            ~Disarm() { sections_arm(false); }
        } disarm;
        sections_arm(true);
        const auto t0 = std::chrono::steady_clock::now();
        if (h->f64) {
            h->a64.init(ctx().stream);
            h->a64.compute(ctx().stream);
        } else {
            h->a32.init(ctx().stream);
            h->a32.compute(ctx().stream);
        }
        SA_HIP(hipStreamSynchronize(ctx().stream));
        const auto t1 = std::chrono::steady_clock::now();
        ms[0] = sections_elapsed(kMarkBuild0, kMarkRefit);
        ms[1] = sections_elapsed(kMarkCompute0, kMarkScans);
        ms[2] = sections_elapsed(kMarkCompress0, kMarkCompress);
        ms[3] = std::chrono::duration<double, std::milli>(t1 - t0).count();
        ms[4] = sections_elapsed(kMarkBuild0, kMarkCompress);
        ms[5] = sections_elapsed(kMarkCompute0, kMarkFill);
        ms[6] = sections_elapsed(kMarkFill, kMarkDelaunayMain);
        ms[7] = sections_elapsed(kMarkDelaunayMain, kMarkDelaunay);
        sections_arm(false);
        check_flags3();
    });
}

}  // extern "C"
