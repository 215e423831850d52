// 2D alpha-shape kernels and the AlphaShape2D C-ABI (see ashape2d.hpp).
#include <chrono>
#include <vector>

#include "ashape2d.hpp"
#include "predicates.hpp"

namespace stream_amd {

uint32_t& ashape2d_edge_capacity() {
    static uint32_t k = 6;
    return k;
}


template <class S> struct Consts2D;
template <> struct Consts2D<float> {
    // 2*delta, delta = sqrt((1+4eps)/(1-3eps)) (DirectAlphaShape2D.cpp:186-188)
    static constexpr float two_delta = 2.0000004172325445139859201780136934301654f;
};
template <> struct Consts2D<double> {
    static constexpr double two_delta = 2.0 + 0x1p-50;  // fl(2 + 7*2^-53)
};

template <class S> struct AshArgs {
    uint32_t n;
    const Sph4<S>* P;          // Morton-ordered points + infinity nodes
    const uint32_t* node;      // LBVH internal nodes
    const BoxD<S, 2>* ibox;    // LBVH internal boxes
    uint32_t* cand;
    uint16_t* tri;
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

// Infinity nodes (cpp:144-163) from the global sphere box.
template <class S> __global__ void infinity_kernel(const BoxD<S, 2>* gbox, Sph4<S>* P, uint32_t n) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const BoxD<S, 2> g = *gbox;
    S avg[2], range[2];
    for (int a = 0; a < 2; ++a) {
        avg[a] = (g.hi[a] + g.lo[a]) * (S)0.5;
        range[a] = g.hi[a] - g.lo[a];
    }
    const S mr = fmax(range[0], range[1]);
    Sph4<S> p;
    p.c[2] = (S)0;
    p.r = (S)0;
    p.c[0] = avg[0] - 2 * mr;
    p.c[1] = avg[1] - 2 * mr;
    P[n] = p;
    p.c[0] = avg[0] + 2 * mr;
    p.c[1] = avg[1] - 2 * mr;
    P[n + 1] = p;
    p.c[0] = avg[0];
    p.c[1] = avg[1] + 2 * mr;
    P[n + 2] = p;
}

// Candidate search (cpp:310-357): stack DFS, right child popped first,
// leaf children tested when their parent is visited.  The child that the
// next pop would return is kept in a register and an internal child's node
// word is loaded with its box; the stack (63 entries + the register = the
// 64-deep stack) holds node words, its first 16 levels in LDS.  The visit
// order, the candidate order and the capacity error are those of the plain
// push/pop loop.
constexpr int kFillLds = 16;  // stack levels held in LDS ([level][thread]); deeper ones in private memory
template <class S> __global__ __launch_bounds__(256) void fill_kernel(AshArgs<S> a) {
    __shared__ uint32_t lstack[kFillLds][256];
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= a.n) return;
    const Sph4<S> si = a.P[tid];
    const S two_delta = Consts2D<S>::two_delta;
    const S rs = two_delta * si.r;
    uint32_t stack[63 - kFillLds];
    int ss = 0;
    bool live = a.n >= 2;
    uint32_t w = live ? a.node[0] : 0u;
    uint32_t nloc = kNInf2D;
    // a valid tree has n-1 internal nodes: the visit cap only guards a
    // malformed topology
    for (uint32_t visits = 0; live; ++visits) {
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
                const BoxD<S, 2> bb = a.ibox[c];
                wc[ch] = a.node[c];
                const S dmaxx = fmax(si.c[0] - bb.hi[0], (S)0);
                const S dmaxy = fmax(si.c[1] - bb.hi[1], (S)0);
                const S dminx = fmax(bb.lo[0] - si.c[0], (S)0);
                const S dminy = fmax(bb.lo[1] - si.c[1], (S)0);
                hit[ch] = dmaxx * dmaxx + dminx * dminx + dmaxy * dmaxy + dminy * dminy < rs * rs;
            } else if (tid != c) {
                const Sph4<S> sj = a.P[c];
                const S ma = fmin(si.r, sj.r);
                const S dx = si.c[0] - sj.c[0];
                const S dy = si.c[1] - sj.c[1];
                const S sq = dx * dx + dy * dy;
                if (sq < (two_delta * ma) * (two_delta * ma)) {
                    a.cand[slab_index(tid, nloc - kNInf2D)] = c;
                    ++nloc;
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
        if (nloc > kMaxFill2D) break;
    }
    a.nineig[tid] = nloc;
}

__device__ __forceinline__ uint16_t pack_tri(uint32_t a, uint32_t b) { return (uint16_t)(a | (b << 8)); }

__device__ __forceinline__ void tag_op(uint64_t& t0, uint64_t& t1, uint64_t& t2, uint64_t& t3, uint32_t v, bool xr) {
    const uint64_t bit = 1ull << (v & 63u);
    const uint32_t w = v >> 6;
    if (xr) {
        t0 ^= (w == 0) ? bit : 0ull;
        t1 ^= (w == 1) ? bit : 0ull;
        t2 ^= (w == 2) ? bit : 0ull;
        t3 ^= (w == 3) ? bit : 0ull;
    } else {
        t0 |= (w == 0) ? bit : 0ull;
        t1 |= (w == 1) ? bit : 0ull;
        t2 |= (w == 2) ? bit : 0ull;
        t3 |= (w == 3) ? bit : 0ull;
    }
}

// Local Delaunay + alpha filter of node si (cpp:373-600).  EXACT=false
// returns false as soon as a filter cannot certify a sign (the node is then
// redone with EXACT=true, which resolves such cases exactly with the
// simplified simulation of simplicity of cpp:444-475).
template <class S, bool EXACT> __device__ bool local_delaunay(const AshArgs<S>& a, uint32_t si, double* ws) {
    const uint32_t n = a.n;
    const Sph4<S>* P = a.P;
    const uint32_t nn = a.nineig[si];
    auto neig = [&](uint32_t j) -> uint32_t { return j < kNInf2D ? n + j : a.cand[slab_index(si, j - kNInf2D)]; };
    auto T = [&](uint32_t ti) -> uint16_t& { return a.tri[slab_index(si, ti)]; };

    T(0) = pack_tri(1, 2);
    T(1) = pack_tri(2, 0);
    T(2) = pack_tri(0, 1);
    for (uint32_t j = 0; j + kNInf2D < nn; ++j) a.active[slab_index(si, j)] = 0;
    if (nn - kNInf2D < 2) {
        a.node_is_bnd[si] = 1;
        a.nelem_out[si] = 0;
        a.nelem[si] = 0;
        a.nfneig[si] = 0;
        return true;
    }
    const S pa[2] = {P[si].c[0], P[si].c[1]};
    uint64_t t0, t1, t2, t3;
    uint32_t tlast = 3;
    bool over = false;  // an exact expansion outgrew the workspace
    for (uint32_t ni = kNInf2D; ni < nn; ++ni) {
        t0 = t1 = t2 = t3 = 0;
        const uint32_t sj = neig(ni);
        const S pd[2] = {P[sj].c[0], P[sj].c[1]};
        uint32_t rsize = 0;
        if constexpr (!EXACT) {
            // Fast path (filters only; a node whose filter cannot decide is
            // redone by the exact kernel, which takes the sequential loop below).
            // Conflict flags of the current triangles with the new point, bit k
            // of word k >> 6 (f0..f3) for position k: every triangle is tested
            // once, two per step so that their loads overlap; the next pair's
            // point ids load while the current pair is tested.
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
            auto pair_ids = [&](uint32_t p0, uint32_t (&o)[2][2]) {
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const uint16_t e = T(u == 0 || p0 + 1 >= tlast ? p0 : p0 + 1);
                    o[u][0] = neig(e & 255u);
                    o[u][1] = neig(e >> 8);
                }
            };
            uint32_t ids[2][2];
            if (tlast) pair_ids(0, ids);
            for (uint32_t p0 = 0; p0 < tlast; p0 += 2) {
                const bool two = p0 + 1 < tlast;
                S pb[2][2], pc[2][2];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const Sph4<S> sb = P[ids[u][0]];
                    const Sph4<S> sc = P[ids[u][1]];
                    pb[u][0] = sb.c[0];
                    pb[u][1] = sb.c[1];
                    pc[u][0] = sc.c[0];
                    pc[u][1] = sc.c[1];
                }
                uint32_t nids[2][2];
                const bool more = p0 + 2 < tlast;
                if (more) pair_ids(p0 + 2, nids);
                S orientation[2], det[2], err[2];
                bool certain[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    orientation[u] = orient2d_filter(pa, pb[u], pc[u], &certain[u]);
                    det[u] = incircle_filter(pa, pb[u], pc[u], pd, &err[u]);
                }
                if (!certain[0] || !((det[0] > err[0]) || (-det[0] > err[0]))) return false;
                if (two && (!certain[1] || !((det[1] > err[1]) || (-det[1] > err[1])))) return false;
                if (orientation[0] < (S)0) det[0] = -det[0];
                if (orientation[1] < (S)0) det[1] = -det[1];
                if (det[0] > (S)0) set_flag(p0, true);
                if (two && det[1] > (S)0) set_flag(p0 + 1, true);
                if (more) {
#pragma unroll
                    for (int u = 0; u < 2; ++u) {
                        ids[u][0] = nids[u][0];
                        ids[u][1] = nids[u][1];
                    }
                }
            }
            // The cavity: the sequential order of cpp:419-495 (a conflicting
            // triangle tags its two edges and is replaced by the last one, which
            // is tested next at the same position) replayed from the flags.
            for (uint32_t ti = next_flag(0); ti < tlast; ti = next_flag(ti)) {
                const uint16_t e = T(ti);
                tag_op(t0, t1, t2, t3, e & 255u, true);
                tag_op(t0, t1, t2, t3, e >> 8, true);
                ++rsize;
                --tlast;
                T(ti) = T(tlast);
                set_flag(ti, flag(tlast));
            }
        } else {
            uint32_t ti = 0;
            while (ti < tlast) {
                const uint16_t e = T(ti);
                const uint32_t ea = e & 255u, eb = e >> 8;
                uint32_t i1 = si, i2 = neig(ea), i3 = neig(eb), i4 = sj;
                const Sph4<S> sb = P[i2];
                const Sph4<S> sc = P[i3];
                S pb[2] = {sb.c[0], sb.c[1]};
                S pc[2] = {sc.c[0], sc.c[1]};
                bool certain;
                S orientation = orient2d_filter(pa, pb, pc, &certain);
                if (!certain) {
                    if constexpr (!EXACT) {
                        return false;
                    } else {
                        // FP32 inputs: the FP64 filter on the exactly widened
                        // coordinates first; the exact stage if it cannot tell
                        bool c64 = false;
                        if constexpr (sizeof(S) != sizeof(double)) {
                            const double qa[2] = {pa[0], pa[1]}, qb[2] = {pb[0], pb[1]}, qc[2] = {pc[0], pc[1]};
                            const double o64 = orient2d_filter(qa, qb, qc, &c64);
                            if (c64) orientation = o64 < 0.0 ? (S)-1 : (S)1;
                        }
                        if (!c64) orientation = (S)orient2d_exact(pa, pb, pc, ws, kExactWsDoubles, &over);
                    }
                }
                S err;
                S det = incircle_filter(pa, pb, pc, pd, &err);
                bool c64 = false;
                if constexpr (EXACT && sizeof(S) != sizeof(double)) {
                    if (!((det > err) || (-det > err))) {
                        const double wa[2] = {pa[0], pa[1]}, wb[2] = {pb[0], pb[1]}, wc[2] = {pc[0], pc[1]},
                                     wd[2] = {pd[0], pd[1]};
                        double e64;
                        const double d64 = incircle_filter(wa, wb, wc, wd, &e64);
                        c64 = (d64 > e64) || (-d64 > e64);
                        if (c64) det = d64 > 0.0 ? (S)1 : (S)-1;
                    }
                }
                if (!c64 && !((det > err) || (-det > err))) {
                    if constexpr (!EXACT) {
                        return false;
                    } else {
                        S qa[2] = {pa[0], pa[1]}, qb[2] = {pb[0], pb[1]}, qc[2] = {pc[0], pc[1]}, qd[2] = {pd[0], pd[1]};
                        bool swap = false;
                        auto cs = [&](uint32_t& x, uint32_t& y, S* px, S* py) {
                            if (x > y) {
                                const uint32_t t = x;
                                x = y;
                                y = t;
                                const S u0 = px[0], u1 = px[1];
                                px[0] = py[0];
                                px[1] = py[1];
                                py[0] = u0;
                                py[1] = u1;
                                swap = !swap;
                            }
                        };
                        cs(i1, i3, qa, qc);
                        cs(i2, i4, qb, qd);
                        cs(i1, i2, qa, qb);
                        cs(i3, i4, qc, qd);
                        cs(i2, i3, qb, qc);
                        det = (S)incircle_exact(qa, qb, qc, qd, ws, kExactWsDoubles, &over);
                        int depth = 0;
                        while (det == (S)0) {
                            ++depth;
                            if (depth == 1) {
                                bool c1;
                                det = orient2d_filter(qb, qc, qd, &c1);
                                if (!c1) det = (S)orient2d_exact(qb, qc, qd, ws, kExactWsDoubles, &over);
                            } else if (depth == 2) {
                                bool c2;
                                det = orient2d_filter(qa, qc, qd, &c2);
                                if (!c2) det = (S)orient2d_exact(qa, qc, qd, ws, kExactWsDoubles, &over);
                                det = -det;
                            } else {
                                det = ((int)swap + (int)(orientation < (S)0)) == 1 ? (S)1 : (S)-1;
                            }
                        }
                        if (swap) det = -det;
                    }
                }
                if (orientation < (S)0) det = -det;
                if (det > (S)0) {
                    tag_op(t0, t1, t2, t3, ea, true);
                    tag_op(t0, t1, t2, t3, eb, true);
                    ++rsize;
                    --tlast;
                    T(ti) = T(tlast);
                } else {
                    ++ti;
                }
            }
        }
        if (rsize) {
            const uint64_t tags[4] = {t0, t1, t2, t3};
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) {
                uint64_t v = tags[j];
                while (v) {
                    const uint32_t r = j * 64 + (uint32_t)__builtin_ctzll(v);
                    v &= v - 1;
                    if (tlast >= kSlabCap) {
                        raise_flag(a.flags, kErrCapacity);
                        break;
                    }
                    T(tlast++) = pack_tri(r, ni);
                }
            }
        }
    }
    // alpha filter (cpp:517-594)
    t0 = t1 = t2 = t3 = 0;
    bool bnd = false;
    uint32_t nout = 0, nloc = 0;
    uint32_t ti = 0;
    while (ti < tlast) {
        const uint16_t e = T(ti);
        const uint32_t ea = e & 255u, eb = e >> 8;
        if (ea < kNInf2D || eb < kNInf2D) {
            bnd = true;
            --tlast;
            T(ti) = T(tlast);
            continue;
        }
        uint32_t i1 = si, i2 = neig(ea), i3 = neig(eb);
        bool swap = false;
        if (i3 < i2) { const uint32_t t = i3; i3 = i2; i2 = t; swap = !swap; }
        if (i2 < i1) { const uint32_t t = i2; i2 = i1; i1 = t; swap = !swap; }
        if (i3 < i2) { const uint32_t t = i3; i3 = i2; i2 = t; swap = !swap; }
        const Sph4<S> p1 = P[i1], p2 = P[i2], p3 = P[i3];
        const S ux = p3.c[0] - p2.c[0], uy = p3.c[1] - p2.c[1];
        const S vx = p1.c[0] - p2.c[0], vy = p1.c[1] - p2.c[1];
        S det = ux * vy - uy * vx;
        if (swap) det = -det;
        const S d12x = p1.c[0] - p2.c[0], d12y = p1.c[1] - p2.c[1];
        const S d23x = p2.c[0] - p3.c[0], d23y = p2.c[1] - p3.c[1];
        const S d31x = p3.c[0] - p1.c[0], d31y = p3.c[1] - p1.c[1];
        const S n12 = d12x * d12x + d12y * d12y;
        const S n23 = d23x * d23x + d23y * d23y;
        const S n31 = d31x * d31x + d31y * d31y;
        const S rsq = ((n12 * n23) * n31) / (((S)4 * det) * det);
        const S aeff = fmin(p3.r, fmin(p1.r, p2.r));
        if (rsq <= aeff * aeff) {
            tag_op(t0, t1, t2, t3, ea, false);
            tag_op(t0, t1, t2, t3, eb, false);
            a.active[slab_index(si, ea - kNInf2D)] += 1;
            a.active[slab_index(si, eb - kNInf2D)] += 1;
            ++nloc;
            if (det < (S)0) T(ti) = pack_tri(eb, ea);
            ++ti;
            nout += (si == i1);
        } else {
            bnd = true;
            --tlast;
            T(ti) = T(tlast);
        }
    }
    a.node_is_bnd[si] = bnd;
    a.nelem_out[si] = (uint8_t)nout;
    a.nelem[si] = (uint8_t)nloc;
    a.nfneig[si] = (uint32_t)(__popcll(t0) + __popcll(t1) + __popcll(t2) + __popcll(t3));
    if (over) raise_flag(a.flags, kErrCapacity);
    return true;
}

template <class S> __global__ __launch_bounds__(256) void delaunay_kernel(AshArgs<S> a) {
    const uint32_t si = blockIdx.x * blockDim.x + threadIdx.x;
    if (si >= a.n) return;
    if (!local_delaunay<S, false>(a, si, nullptr)) {
        const uint32_t k = atomicAdd(&a.queue[0], 1u);
        a.queue[1 + k] = si;
    }
}

// One queued node per wavefront (lane 0; the block is one wavefront): the
// exact evaluations of different nodes happen at different points of their
// insertions, so nodes sharing a wavefront would run their exact stages one
// after another.  The expansion workspace lives in LDS.
template <class S> __global__ __launch_bounds__(64) void delaunay_exact_kernel(AshArgs<S> a) {
    __shared__ double ws[kExactWsDoubles];
    if (threadIdx.x != 0) return;
    const uint32_t cnt = a.queue[0];
    for (uint32_t k = blockIdx.x; k < cnt; k += gridDim.x) local_delaunay<S, true>(a, a.queue[1 + k], ws);
}

// Slab -> CSR (cpp:642-689) with capacity checks.
template <class S>
__global__ __launch_bounds__(256) void compress_kernel(AshArgs<S> a, const uint32_t* __restrict__ row,
                                                       const uint32_t* __restrict__ elemrow, uint32_t* __restrict__ neig,
                                                       uint8_t* __restrict__ edge_is_bnd, uint32_t* __restrict__ elem,
                                                       uint32_t cap_edges, uint32_t cap_elems) {
    const uint32_t si = blockIdx.x * blockDim.x + threadIdx.x;
    if (si >= a.n) return;
    const uint32_t n = a.n;
    auto gneig = [&](uint32_t j) -> uint32_t { return j < kNInf2D ? n + j : a.cand[slab_index(si, j - kNInf2D)]; };
    uint32_t put = row[si];
    const uint32_t nn = a.nineig[si];
    for (uint32_t i = kNInf2D; i < nn; ++i) {
        const uint8_t ec = a.active[slab_index(si, i - kNInf2D)];
        if (ec) {
            if (put < cap_edges) {
                edge_is_bnd[put] = (ec == 1);
                neig[put] = gneig(i);
            } else {
                raise_flag(a.flags, kErrCapacity);
            }
            ++put;
        }
    }
    uint32_t t = 0;
    const uint32_t kend = elemrow[si + 1];
    const uint32_t tcount = a.nelem[si];
    for (uint32_t k = elemrow[si]; k < kend; ++k) {
        uint32_t b = 0, c = 0;
        while (t < tcount) {
            const uint16_t e = a.tri[slab_index(si, t)];
            b = gneig(e & 255u);
            c = gneig(e >> 8);
            if (b >= si && c >= si) break;
            ++t;
        }
        if (t >= tcount) {
            raise_flag(a.flags, kErrCapacity);
            break;
        }
        if (k < cap_elems) {
            elem[3ull * k] = si;
            elem[3ull * k + 1] = b;
            elem[3ull * k + 2] = c;
        } else {
            raise_flag(a.flags, kErrCapacity);
        }
        ++t;
    }
}

template <class S> AshArgs<S> make_args(Ash2D<S>& A) {
    AshArgs<S> a;
    a.n = A.n;
    a.P = A.lbvh.obj_m.p;
    a.node = A.lbvh.node.p;
    a.ibox = A.lbvh.ibox.p;
    a.cand = A.cand.p;
    a.tri = A.tri.p;
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

template <class S> void Ash2D<S>::set_nodes_host(uint32_t n_, const S* xy, const S* alpha) {
    std::vector<Sph4<S>> h(n_);
    for (uint32_t i = 0; i < n_; ++i) {
        h[i].c[0] = xy[2ull * i];
        h[i].c[1] = xy[2ull * i + 1];
        h[i].c[2] = (S)0;
        h[i].r = alpha[i];
    }
    n = n_;
    coords.resize(n ? n : 1);
    h2d(coords.p, h.data(), n);
    inited = computed = false;
}

template <class S> void Ash2D<S>::set_nodes_device(uint32_t n_, const Sph4<S>* d) {
    n = n_;
    coords.resize(n ? n : 1);
    d2d(coords.p, d, n);
    inited = computed = false;
}

template <class S> void Ash2D<S>::init(hipStream_t s) {
    if (n < 3) throw std::invalid_argument("AlphaShape2D needs at least 3 nodes");
    lbvh.build(coords.p, n, bits, kLeafCentre, kNInf2D, s);
    hipLaunchKernelGGL((infinity_kernel<S>), dim3(1), dim3(64), 0, s, lbvh.gbox.p, lbvh.obj_m.p, n);
    const size_t sw = slab_words(n);
    cand.resize(sw);
    tri.resize(sw);
    active.resize(sw);
    nineig.resize(n);
    nfneig.resize(n);
    node_is_bnd.resize(n);
    nelem.resize(n);
    nelem_out.resize(n);
    row.resize(n + 1);
    elemrow.resize(n + 1);
    cap_edges = ashape2d_edge_capacity() * n + 64u;
    cap_elems = 2u * n + 64u;
    neig.resize(cap_edges);
    edge_is_bnd.resize(cap_edges);
    elemglob.resize(3ull * cap_elems);
    queue.resize((size_t)n + 1);
    scan_ws.resize(scan_workspace_words(n));
    inited = true;
    computed = false;
}

template <class S> void Ash2D<S>::compute(hipStream_t s) {
    if (!inited) throw std::invalid_argument("AlphaShape2D_compute called before init");
    AshArgs<S> a = make_args(*this);
    section_mark(s, kMarkCompute0);
    fill<uint32_t>(queue.p, 1, 0u, s);
    hipLaunchKernelGGL((fill_kernel<S>), dim3(grid_for(n, 256)), dim3(256), 0, s, a);
    section_mark(s, kMarkFill);
    hipLaunchKernelGGL((delaunay_kernel<S>), dim3(grid_for(n, 256)), dim3(256), 0, s, a);
    section_mark(s, kMarkDelaunayMain);
    hipLaunchKernelGGL((delaunay_exact_kernel<S>), dim3(kExactWaves), dim3(64), 0, s, a);
    section_mark(s, kMarkDelaunay);
    exclusive_scan_u8(nelem_out.p, elemrow.p, n, scan_ws.p, s);
    exclusive_scan_u32(nfneig.p, row.p, n, scan_ws.p, s);
    section_mark(s, kMarkScans);
    section_mark(s, kMarkCompress0);
    hipLaunchKernelGGL((compress_kernel<S>), dim3(grid_for(n, 256)), dim3(256), 0, s, a, row.p, elemrow.p, neig.p,
                       edge_is_bnd.p, elemglob.p, cap_edges, cap_elems);
    section_mark(s, kMarkCompress);
    computed = true;
}

template <class S> uint32_t Ash2D<S>::n_elems() { return computed ? read_scalar(elemrow.p + n) : 0u; }
template <class S> uint32_t Ash2D<S>::n_edges() { return computed ? read_scalar(row.p + n) : 0u; }
template <class S> uint32_t Ash2D<S>::n_exact_nodes() { return computed ? read_scalar(queue.p) : 0u; }

template struct Ash2D<float>;
template struct Ash2D<double>;

}  // namespace stream_amd

// ---------------------------------------------------------------------------
// C-ABI: the reference's AlphaShape2D_* surface (DirectAlphaShape2D.hpp:48-67,
// DirectAlphaShape2D.cpp:698-753) plus FP64 and query extensions.
// ---------------------------------------------------------------------------
using namespace stream_amd;

struct AlphaShape2D {
    int bits = 32;
    bool f64 = false;
    Ash2D<float> a32;
    Ash2D<double> a64;
    int32_t status = 0;
};

namespace {
template <class F> void run_void(AlphaShape2D* h, F&& f) {
    if (!h) {
        set_last_error("null AlphaShape2D handle");
        return;
    }
    h->status = guarded(f);
}
template <class F> int32_t run_status(AlphaShape2D* h, F&& f) {
    if (!h) {
        set_last_error("null AlphaShape2D handle");
        return kErrArgument;
    }
    h->status = guarded(f);
    return h->status;
}
// Device flag check after a synchronising call.
void check_flags() {
    const uint32_t f = read_scalar(ctx().d_flags + kFlagError);
    if (f) {
        fill<uint32_t>(ctx().d_flags, 1, 0u, ctx().stream);
        throw std::runtime_error(f == kErrCapacity ? "device capacity overflow (neighbour stack / output bound)"
                                                   : "device error flag raised");
    }
}
}  // namespace

extern "C" {

AlphaShape2D* AlphaShape2D_create(void) {
    try {
        (void)ctx();
        return new AlphaShape2D();
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return nullptr;
    }
}

void AlphaShape2D_destroy(AlphaShape2D* h) { delete h; }

// Select the Morton resolution (bits per axis, 1..32) used by the next init.
int32_t AlphaShape2D_set_morton_bits(AlphaShape2D* h, int32_t bits) {
    return run_status(h, [&] {
        if (bits < 1 || bits > 32) throw std::invalid_argument("Morton bits per axis must be in [1, 32]");
        h->bits = bits;
    });
}

// FP32 nodes: coords row-major [2N], alpha [N] (reference signature).
void AlphaShape2D_set_nodes(AlphaShape2D* h, uint32_t nnodes, const float* coords, const float* alpha) {
    run_void(h, [&] {
        h->f64 = false;
        h->a32.set_nodes_host(nnodes, coords, alpha);
    });
}

// FP64 nodes: coords row-major [2N], alpha [N].
int32_t AlphaShape2D_set_nodes_f64(AlphaShape2D* h, uint32_t nnodes, const double* coords, const double* alpha) {
    return run_status(h, [&] {
        h->f64 = true;
        h->a64.set_nodes_host(nnodes, coords, alpha);
    });
}

// FP64 nodes already in device memory, as rows (x, y, pad, alpha) of 4 doubles.
int32_t AlphaShape2D_set_nodes_device_f64(AlphaShape2D* h, uint32_t nnodes, const double* d_xypr) {
    return run_status(h, [&] {
        h->f64 = true;
        h->a64.set_nodes_device(nnodes, reinterpret_cast<const Sph4<double>*>(d_xypr));
    });
}

void AlphaShape2D_init(AlphaShape2D* h) {
    run_void(h, [&] {
        if (h->f64) {
            h->a64.bits = h->bits;
            h->a64.init(ctx().stream);
        } else {
            h->a32.bits = h->bits;
            h->a32.init(ctx().stream);
        }
    });
}

// compute + compress (reference AlphaShape2D_compute, cpp:729-732)
void AlphaShape2D_compute(AlphaShape2D* h) {
    run_void(h, [&] {
        if (h->f64)
            h->a64.compute(ctx().stream);
        else
            h->a32.compute(ctx().stream);
    });
}

uint32_t AlphaShape2D_get_nelem(const AlphaShape2D* hc) {
    AlphaShape2D* h = const_cast<AlphaShape2D*>(hc);
    uint32_t r = 0;
    run_void(h, [&] {
        r = h->f64 ? h->a64.n_elems() : h->a32.n_elems();
        check_flags();
    });
    return r;
}

// Triangles [3*nelem], Morton ids, each owned by its lowest id.
void AlphaShape2D_get_elem(const AlphaShape2D* hc, uint32_t* elems) {
    AlphaShape2D* h = const_cast<AlphaShape2D*>(hc);
    run_void(h, [&] {
        const uint32_t ne = h->f64 ? h->a64.n_elems() : h->a32.n_elems();
        check_flags();
        const uint32_t cap = h->f64 ? h->a64.cap_elems : h->a32.cap_elems;
        if (ne > cap) throw std::runtime_error("triangle count exceeds the output capacity");
        d2h(elems, h->f64 ? h->a64.elemglob.p : h->a32.elemglob.p, 3ull * ne);
    });
}

// Morton-ordered node coordinates [2N] (FP32 handle).
void AlphaShape2D_get_ordered_nodes(AlphaShape2D* h, float* nodes) {
    run_void(h, [&] {
        if (h->f64) throw std::invalid_argument("handle holds FP64 nodes: use AlphaShape2D_get_ordered_nodes_f64");
        std::vector<Sph4<float>> v(h->a32.n);
        d2h(v.data(), h->a32.lbvh.obj_m.p, h->a32.n);
        for (uint32_t i = 0; i < h->a32.n; ++i) {
            nodes[2 * i] = v[i].c[0];
            nodes[2 * i + 1] = v[i].c[1];
        }
    });
}

int32_t AlphaShape2D_get_ordered_nodes_f64(AlphaShape2D* h, double* nodes) {
    return run_status(h, [&] {
        if (!h->f64) throw std::invalid_argument("handle holds FP32 nodes");
        std::vector<Sph4<double>> v(h->a64.n);
        d2h(v.data(), h->a64.lbvh.obj_m.p, h->a64.n);
        for (uint32_t i = 0; i < h->a64.n; ++i) {
            nodes[2 * i] = v[i].c[0];
            nodes[2 * i + 1] = v[i].c[1];
        }
    });
}

// Number of directed alpha-shape edges (sum of final neighbour counts).
uint32_t AlphaShape2D_get_nedge(const AlphaShape2D* hc) {
    AlphaShape2D* h = const_cast<AlphaShape2D*>(hc);
    uint32_t r = 0;
    run_void(h, [&] { r = h->f64 ? h->a64.n_edges() : h->a32.n_edges(); });
    return r;
}

// Edge CSR in Morton ids (AlphaShape2D::getEdge, cpp:90-98) and the
// free-surface flags of edges (count == 1) and nodes (getBoundaryNodes,
// cpp:101-105).  Any pointer may be null.
int32_t AlphaShape2D_get_edges(const AlphaShape2D* hc, uint32_t* row, uint32_t* neig, uint8_t* edge_is_bnd,
                               uint8_t* node_is_bnd) {
    AlphaShape2D* h = const_cast<AlphaShape2D*>(hc);
    return run_status(h, [&] {
        auto go = [&](auto& A) {
            const uint32_t ne = A.n_edges();
            check_flags();
            if (ne > A.cap_edges) throw std::runtime_error("edge count exceeds the output capacity");
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

// perm[i] = input index of Morton slot i (AlphaShape2D::getPermutation,
// cpp:75-80; x_i = x_r[perm^-1[i]]).
int32_t AlphaShape2D_get_permutation(const AlphaShape2D* hc, uint32_t* perm) {
    AlphaShape2D* h = const_cast<AlphaShape2D*>(hc);
    return run_status(h, [&] {
        if (h->f64)
            d2h(perm, h->a64.lbvh.perm, h->a64.n);
        else
            d2h(perm, h->a32.lbvh.perm, h->a32.n);
    });
}

// Nodes whose local triangulation needed the exact predicate stage.
uint32_t AlphaShape2D_get_nexact(const AlphaShape2D* hc) {
    AlphaShape2D* h = const_cast<AlphaShape2D*>(hc);
    uint32_t r = 0;
    run_void(h, [&] { r = h->f64 ? h->a64.n_exact_nodes() : h->a32.n_exact_nodes(); });
    return r;
}

// Status of the last call on this handle (0 = ok); the message is in
// Stream_last_error().
int32_t AlphaShape2D_last_status(const AlphaShape2D* h) { return h ? h->status : kErrArgument; }


// Nodal fields between input order and Morton order (permuteForward /
// permuteBackward, include/mesh/DirectAlphaShape2D.hpp): forward out[i] =
// in[perm[i]] gives the field in the Morton order of the outputs, backward
// out[perm[i]] = in[i] returns it to input order.  n elements of
// elem_bytes each; host arrays (copied) or device buffers (_device, on the
// library stream, not synchronised).  Needs init (the permutation).
}  // extern "C"

namespace {
template <class F> int32_t permute_call(const AlphaShape2D* hc, F&& f) {
    AlphaShape2D* h = const_cast<AlphaShape2D*>(hc);
    return run_status(h, [&] {
        const bool ok = h->f64 ? h->a64.inited : h->a32.inited;
        if (!ok) throw std::invalid_argument("AlphaShape2D: permutation needs init (set_nodes)");
        const uint32_t* perm = h->f64 ? h->a64.lbvh.perm : h->a32.lbvh.perm;
        const uint32_t n = h->f64 ? h->a64.n : h->a32.n;
        f(perm, n);
    });
}
}  // namespace

extern "C" {

int32_t AlphaShape2D_permute_forward(const AlphaShape2D* h, const void* in, void* out, uint32_t elem_bytes) {
    return permute_call(h, [&](const uint32_t* perm, uint32_t n) {
        permute_host_elements(in, out, perm, n, elem_bytes, true);
    });
}
int32_t AlphaShape2D_permute_backward(const AlphaShape2D* h, const void* in, void* out, uint32_t elem_bytes) {
    return permute_call(h, [&](const uint32_t* perm, uint32_t n) {
        permute_host_elements(in, out, perm, n, elem_bytes, false);
    });
}
int32_t AlphaShape2D_permute_forward_device(const AlphaShape2D* h, const void* d_in, void* d_out, uint32_t elem_bytes) {
    return permute_call(h, [&](const uint32_t* perm, uint32_t n) {
        permute_elements(d_in, d_out, perm, n, elem_bytes, true, ctx().stream);
    });
}
int32_t AlphaShape2D_permute_backward_device(const AlphaShape2D* h, const void* d_in, void* d_out, uint32_t elem_bytes) {
    return permute_call(h, [&](const uint32_t* perm, uint32_t n) {
        permute_elements(d_in, d_out, perm, n, elem_bytes, false, ctx().stream);
    });
}


// One init + compute timed as testcases/IMR26/uniform2d.cpp times them:
// ms[0] tlbvh (LBVH build), ms[1] tashape (candidate search, local
// Delaunay + filter, element/edge scans), ms[2] tcompress (compaction to
// CSR), ms[3] the host wall time of both calls, ms[4] the GPU time from the
// build's start to the compaction's end; within tashape, ms[5] the
// candidate search, ms[6] the local Delaunay + filter of every node, ms[7]
// the exact redo of the nodes whose filters could not decide a sign.  GPU times come from HIP events
// on the library stream between the launches.
int32_t AlphaShape2D_timed_run(AlphaShape2D* h, double* ms) {
    return run_status(h, [&] {
        if (!ms) throw std::invalid_argument("AlphaShape2D_timed_run: null ms");
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
        check_flags();
    });
}

}  // extern "C"
