// Full 2D Delaunay triangulation by per-node local stars.  Replaces
// src/mesh/mesh.cpp:45-862 and include/mesh/mesh.hpp:84-332 (Mesh<2>,
// fp_tt = float).
//
// Every node owns a star: triangles (node, neighbour a, neighbour b) over
// local neighbour indices (0..2 are the super-triangle vertices, 3.. the
// inserted neighbours).  All phases act on one node's star and read only
// the coordinates and the agglomerative BVH, so each node is one thread and
// no thread waits on another.  The reference runs the circumcircle search
// as up to 32 rounds of (find kernel, compaction, insert kernel); the rounds
// of a node do not depend on other nodes, so here one kernel runs all of a
// node's rounds back to back and the star never leaves the thread between
// rounds except through its own cache lines.
//
// Star layout, for the lanes of a wave to touch consecutive words: the
// stars of nodes 64b..64b+63 form one block; slot j of lane l sits at
// (b * kCap + j) * 64 + l, for the triangles (u16: a | b << 8) and the
// neighbours (u32, local index j >= 3 at slot j - 3).
//
// Capacity: 64 triangles, local indices below 64 (64-bit cavity set) and a
// 64-entry search heap per node.  The reference has 32 of each with no
// checks (past them its result is undefined); below its limits every step is
// identical.  A node passing 64 stops changing and is counted.
#include <vector>

#include "lbvha.hpp"
#include "predicates.hpp"

namespace stream_amd {

namespace {

constexpr uint32_t kM2Inf = 3;      // mesh.hpp:91
constexpr uint32_t kM2Cap = 64;     // triangles and local indices per node
constexpr uint32_t kM2Rounds = 32;  // cur_max_nneig rounds (mesh.cpp:456)
constexpr uint32_t kM2Morton = 10;  // n_init_insert (mesh.hpp:104)
constexpr int kM2Heap = 64;
constexpr int kM2Stack = 32;        // TraversalStack<u32, 32> (mesh.cpp:43)
constexpr uint32_t kM2Block = 64;

enum M2Stat : uint32_t { kM2Overflow = 0, kM2MaxRounds = 1, kM2Exact = 2, kM2StatWords = 4 };

struct M2Star {
    uint16_t* e;
    uint32_t* nb;
    __device__ __forceinline__ uint16_t elem(uint32_t j) const { return e[(size_t)j * kM2Block]; }
    __device__ __forceinline__ void set_elem(uint32_t j, uint16_t v) const { e[(size_t)j * kM2Block] = v; }
    __device__ __forceinline__ uint32_t glob(uint32_t j, uint32_t n) const {
        return j >= kM2Inf ? nb[(size_t)(j - kM2Inf) * kM2Block] : n + j;
    }
    __device__ __forceinline__ void set_neig(uint32_t j, uint32_t g) const { nb[(size_t)(j - kM2Inf) * kM2Block] = g; }
};

__device__ __forceinline__ M2Star m2_star(uint16_t* elem, uint32_t* neig, uint32_t t) {
    const size_t base = (size_t)(t / kM2Block) * kM2Cap * kM2Block + (t % kM2Block);
    return {elem + base, neig + base};
}

// Robust orient2d on float nodes: filter, exact sign when uncertain.
// Float inputs: products of two floats give at most 12 expansion
// components, far below the 64-double workspace.
__device__ __noinline__ float m2_orient_exact(const float* a, const float* b, const float* c) {
    double ws[64];
    return (float)orient2d_exact(a, b, c, ws, 64);
}
__device__ __forceinline__ float m2_orient(const float* a, const float* b, const float* c) {
    bool certain;
    const float d = orient2d_filter(a, b, c, &certain);
    return certain ? d : m2_orient_exact(a, b, c);
}

// Uncertain incircle: exact sign after sorting the four indices, then the
// SoS fallbacks (mesh.hpp:303-325).  Degree-4 monomials of float inputs span
// at most ~1100 bits, i.e. at most 22 non-overlapping components: 64
// doubles of workspace bound every intermediate expansion.
__device__ __noinline__ float m2_incircle_sos_exact(uint32_t i1, uint32_t i2, uint32_t i3, uint32_t i4,
                                                   const float2* __restrict__ P, float orientation) {
    double ws[64];
    float qa[2] = {P[i1].x, P[i1].y}, qb[2] = {P[i2].x, P[i2].y};
    float qc[2] = {P[i3].x, P[i3].y}, qd[2] = {P[i4].x, P[i4].y};
    bool swap = false;
    auto cs = [&](uint32_t& x, uint32_t& y, float* px, float* py) {
        if (x > y) {
            const uint32_t t = x;
            x = y;
            y = t;
            const float u0 = px[0], u1 = px[1];
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
    float det = (float)incircle_exact(qa, qb, qc, qd, ws, 64);
    int depth = 0;
    while (det == 0.0f) {
        ++depth;
        if (depth == 1)
            det = m2_orient(qb, qc, qd);
        else if (depth == 2)
            det = -m2_orient(qa, qc, qd);
        else
            det = ((int)swap + (int)(orientation < 0.0f)) == 1 ? 1.0f : -1.0f;
    }
    return swap ? -det : det;
}

// incircle_SoS (mesh.hpp:287-329): > 0 when node i4 is inside the
// circumcircle of (i1, i2, i3), symbolic perturbation on ties.
__device__ __forceinline__ float m2_incircle_sos(uint32_t i1, uint32_t i2, uint32_t i3, uint32_t i4,
                                                 const float2* __restrict__ P, uint32_t* stats) {
    const float pa[2] = {P[i1].x, P[i1].y}, pb[2] = {P[i2].x, P[i2].y};
    const float pc[2] = {P[i3].x, P[i3].y}, pd[2] = {P[i4].x, P[i4].y};
    const float orientation = m2_orient(pa, pb, pc);
    float err;
    float det = incircle_filter(pa, pb, pc, pd, &err);
    if (!((det > err) || (-det > err))) {
        atomicAdd(&stats[kM2Exact], 1u);
        det = m2_incircle_sos_exact(i1, i2, i3, i4, P, orientation);
    }
    return orientation < 0.0f ? -det : det;
}

struct M2Args {
    uint32_t n;
    const float2* P;  // n Morton-ordered nodes + 3 super nodes
    uint16_t* elem;
    uint32_t* neig;
    uint32_t* nelem;
    uint32_t* nlocal;  // bit 31: frozen
    uint32_t* stats;
};

constexpr uint32_t kFrozen = 1u << 31;

// Bowyer-Watson insertion of node g into the star over triangles
// [first, T) (mesh.cpp:210-245, 359-395, 639-675).  Returns false when the
// star overflows; the star is then frozen (its contents are unspecified).
__device__ bool m2_insert(const M2Args& A, const M2Star& st, uint32_t tid, uint32_t g, uint32_t first, uint32_t& T,
                          uint32_t& nloc) {
    const uint32_t loc = nloc;
    if (loc - kM2Inf < kM2Cap) st.set_neig(loc, g);
    uint64_t shell = 0;
    uint32_t ti = first;
    while (ti < T) {
        const uint16_t el = st.elem(ti);
        const uint32_t ea = el & 255u, eb = el >> 8;
        if (m2_incircle_sos(tid, st.glob(ea, A.n), st.glob(eb, A.n), g, A.P, A.stats) > 0.0f) {
            shell ^= 1ull << ea;
            shell ^= 1ull << eb;
            --T;
            st.set_elem(ti, st.elem(T));
        } else {
            ++ti;
        }
    }
    if (shell) {
        if (loc >= kM2Cap || T + 2 > kM2Cap) return false;
        st.set_elem(T++, (uint16_t)(__ffsll((long long)shell) - 1) | (uint16_t)(loc << 8));
        st.set_elem(T++, (uint16_t)(63 - __clzll((long long)shell)) | (uint16_t)(loc << 8));
        nloc = loc + 1;
    }
    return true;
}

__device__ __forceinline__ void m2_freeze(const M2Args& A, uint32_t tid, uint32_t T, uint32_t nloc) {
    A.nelem[tid] = T;
    A.nlocal[tid] = nloc | kFrozen;
    atomicAdd(&A.stats[kM2Overflow], 1u);
}

// Mesh::init (mesh.cpp:68-165): Morton-ordered nodes, super-triangle at
// avg +- (max extent + 10) and the 3-triangle stars.
__global__ __launch_bounds__(256) void m2_init_kernel(uint32_t n, const Sph4<float>* __restrict__ obj_m,
                                                      const BoxD<float, 2>* __restrict__ gbox, float2* __restrict__ P,
                                                      uint16_t* elem, uint32_t* neig, uint32_t* nelem,
                                                      uint32_t* nlocal, uint32_t* nelem_out) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n) return;
    P[tid] = make_float2(obj_m[tid].c[0], obj_m[tid].c[1]);
    if (tid == 0) {
        const BoxD<float, 2> g = *gbox;
        const float avg0 = 0.5f * (g.hi[0] + g.lo[0]), avg1 = 0.5f * (g.hi[1] + g.lo[1]);
        const float mr = fmaxf(g.hi[0] - g.lo[0], g.hi[1] - g.lo[1]) + 10.0f;
        P[n] = make_float2(avg0 - mr, avg1 - mr);
        P[n + 1] = make_float2(avg0 + mr, avg1 - mr);
        P[n + 2] = make_float2(avg0, avg1 + mr);
        nelem_out[0] = 0;
    }
    const M2Star st = m2_star(elem, neig, tid);
    st.set_elem(0, 1 | (2 << 8));
    st.set_elem(1, 2 | (0 << 8));
    st.set_elem(2, 0 | (1 << 8));
    nelem[tid] = 3;
    nlocal[tid] = kM2Inf;
    nelem_out[tid + 1] = 0;
}

// Mesh::insert_morton_neighbors (mesh.cpp:167-252): the nodes at Morton
// distance <= 5 (10 at the ends).
__global__ __launch_bounds__(kM2Block) void m2_morton_kernel(M2Args A) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = A.n;
    if (tid >= n) return;
    uint32_t nloc = A.nlocal[tid];
    if (nloc & kFrozen) return;
    uint32_t T = A.nelem[tid];
    const M2Star st = m2_star(A.elem, A.neig, tid);
    const uint32_t h = kM2Morton / 2;
    uint32_t lo = tid - h, hi = tid + h;
    if (tid < h) {
        lo = 0;
        hi = n < kM2Morton ? n : kM2Morton + 1;
    }
    if (tid + h >= n) {
        lo = n < kM2Morton ? 0 : n - kM2Morton;
        hi = n;
    }
    for (uint32_t g = lo; g < hi; ++g) {
        if (g == tid) continue;
        if (!m2_insert(A, st, tid, g, 0, T, nloc)) {
            m2_freeze(A, tid, T, nloc);
            return;
        }
    }
    A.nelem[tid] = T;
    A.nlocal[tid] = nloc;
}

// Cells of the implicit kd-tree of the BVH splits (mesh.cpp:274-311):
// descend towards leaf tid, cutting the global box at the mid-plane between
// the two children's boxes on the axis of the highest differing Morton bit.
__global__ __launch_bounds__(256) void m2_cells_kernel(uint32_t n, const uint32_t* __restrict__ root,
                                                       const uint32_t* __restrict__ children,
                                                       const uint32_t* __restrict__ split,
                                                       const uint32_t* __restrict__ range,
                                                       const uint64_t* __restrict__ codes,
                                                       const BoxD<float, 2>* __restrict__ data,
                                                       const BoxD<float, 2>* __restrict__ gbox,
                                                       BoxD<float, 2>* __restrict__ cell, uint32_t* flags) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n) return;
    BoxD<float, 2> bb = *gbox;
    uint32_t cur = *root;
    if (tid == 0) cell[cur] = bb;  // never read; the reference leaves it unset
    for (uint32_t level = 0; level < 2 * n; ++level) {
        const uint32_t k = cur - n;
        const uint32_t sp = split[k];
        const uint64_t x = codes[sp] ^ codes[sp + 1];
        const int bit = x ? __clzll((long long)x) : 64;
        const int axis = (bit + 1) % 2;
        const uint32_t c0 = children[2 * k], c1 = children[2 * k + 1];
        const float coord = 0.5f * (data[c0].hi[axis] + data[c1].lo[axis]);
        const uint32_t child = tid > sp ? c1 : c0;
        if (child == c0)
            bb.hi[axis] = coord;
        else
            bb.lo[axis] = coord;
        if (child < n || range[2 * (child - n)] == tid) cell[child] = bb;
        if (child < n) return;
        cur = child;
    }
    raise_flag(flags, kErrCapacity);
}

// Mesh::insert_BVH_neighbors (mesh.cpp:313-413): every leaf whose cell
// touches the node's cell, in stack order; an internal node met with a full
// stack has its whole leaf range inserted.
__global__ __launch_bounds__(kM2Block) void m2_bvh_kernel(M2Args A, const uint32_t* __restrict__ root,
                                                          const uint32_t* __restrict__ children,
                                                          const uint32_t* __restrict__ range,
                                                          const BoxD<float, 2>* __restrict__ cell) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = A.n;
    if (tid >= n) return;
    uint32_t nloc = A.nlocal[tid];
    if (nloc & kFrozen) return;
    uint32_t T = A.nelem[tid];
    const M2Star st = m2_star(A.elem, A.neig, tid);
    const BoxD<float, 2> bl = cell[tid];
    uint32_t stack[kM2Stack];
    int len = 0;
    stack[len++] = *root;
    while (len) {
        const uint32_t cur = stack[--len];
        for (int ch = 0; ch < 2; ++ch) {
            const uint32_t child = children[2 * (cur - n) + ch];
            const BoxD<float, 2> bn = cell[child];
            bool nb = true;
            nb &= !(bl.hi[0] < bn.lo[0] || bl.lo[0] > bn.hi[0]);
            nb &= !(bl.hi[1] < bn.lo[1] || bl.lo[1] > bn.hi[1]);
            if (!nb) continue;
            if (child >= n && len < kM2Stack) {
                stack[len++] = child;
                continue;
            }
            const uint32_t rmin = child < n ? child : range[2 * (child - n)];
            const uint32_t rmax = child < n ? child : range[2 * (child - n) + 1];
            for (uint32_t o = rmin; o <= rmax; ++o) {
                if (o == tid) continue;
                if (!m2_insert(A, st, tid, o, 0, T, nloc)) {
                    m2_freeze(A, tid, T, nloc);
                    return;
                }
            }
        }
    }
    A.nelem[tid] = T;
    A.nlocal[tid] = nloc;
}

// Min-heap with the reference's sift rules (traversal_stack.hpp:31-90).
struct M2Heap {
    uint32_t idx[kM2Heap];
    float cost[kM2Heap];
    int len = 0;
    __device__ __forceinline__ bool push(uint32_t i, float c) {
        if (len >= kM2Heap) return false;
        int k = len++;
        idx[k] = i;
        cost[k] = c;
        int p = (k - 1) >> 1;
        while (k > 0 && cost[p] > cost[k]) {
            const uint32_t ti = idx[k];
            const float tc = cost[k];
            idx[k] = idx[p];
            cost[k] = cost[p];
            idx[p] = ti;
            cost[p] = tc;
            k = p;
            p = (k - 1) >> 1;
        }
        return true;
    }
    __device__ __forceinline__ void pop() {
        --len;
        idx[0] = idx[len];
        cost[0] = cost[len];
        int k = 0;
        while (true) {
            const int l = 2 * k + 1, r = 2 * k + 2;
            int s = k;
            if (l < len && cost[l] < cost[s]) s = l;
            if (r < len && cost[r] < cost[s]) s = r;
            if (s == k) break;
            const uint32_t ti = idx[k];
            const float tc = cost[k];
            idx[k] = idx[s];
            cost[k] = cost[s];
            idx[s] = ti;
            cost[s] = tc;
            k = s;
        }
    }
};

// Nearest node to tid strictly inside the circumcircle of one of the star's
// triangles from nds on, by a best-first BVH search bounded by the
// circumcircle (mesh.cpp:465-566).  Triangles found empty move to the
// front.  The heap is shared by the triangles of one call.  Returns tid when
// every triangle is empty, 0xffffffff on heap overflow.
__device__ uint32_t m2_find(const M2Args& A, const M2Star& st, uint32_t tid, uint32_t T, uint32_t& nds,
                            uint32_t root, const uint32_t* __restrict__ children,
                            const BoxD<float, 2>* __restrict__ data) {
    const uint32_t n = A.n;
    M2Heap heap;
    uint32_t g = tid;
    const float2 p1 = A.P[tid];
    for (uint32_t et = nds; et < T; ++et) {
        const uint16_t el = st.elem(et);
        const uint32_t nb0 = st.glob(el & 255u, n), nb1 = st.glob(el >> 8, n);
        g = tid;
        const float2 p2 = A.P[nb0];
        const float2 p3 = A.P[nb1];
        const float bax = p2.x - p1.x, bay = p2.y - p1.y;
        const float cax = p3.x - p1.x, cay = p3.y - p1.y;
        const float cbx = p3.x - p2.x, cby = p3.y - p2.y;
        const float l1 = bax * bax + bay * bay;
        const float l2 = cax * cax + cay * cay;
        const float l3 = cbx * cbx + cby * cby;
        const float det = bax * cay - bay * cax;
        const float rsq = 1.0000001f * ((l1 * l2 * l3) / (4.f * det * det));
        const float ox = p3.x - 0.5f * (l2 * cby - l3 * cay) / det;
        const float oy = p3.y - 0.5f * (l3 * cax - l2 * cbx) / det;
        if (!heap.push(root, 0.0f)) return 0xffffffffu;
        float best = 4.f * rsq;
        while (heap.len) {
            if (heap.cost[0] >= best) break;
            const uint32_t cur = heap.idx[0];
            heap.pop();
            for (int ch = 0; ch < 2; ++ch) {
                const uint32_t child = children[2 * (cur - n) + ch];
                if (child >= n) {
                    const BoxD<float, 2> b = data[child];
                    float dmaxx = fmaxf(p1.x - b.hi[0], 0.0f);
                    float dmaxy = fmaxf(p1.y - b.hi[1], 0.0f);
                    float dminx = fmaxf(b.lo[0] - p1.x, 0.0f);
                    float dminy = fmaxf(b.lo[1] - p1.y, 0.0f);
                    const float sqp = dmaxx * dmaxx + dminx * dminx + dmaxy * dmaxy + dminy * dminy;
                    if (sqp >= best) continue;
                    dmaxx = fmaxf(ox - b.hi[0], 0.0f);
                    dmaxy = fmaxf(oy - b.hi[1], 0.0f);
                    dminx = fmaxf(b.lo[0] - ox, 0.0f);
                    dminy = fmaxf(b.lo[1] - oy, 0.0f);
                    const float sqd = dmaxx * dmaxx + dminx * dminx + dmaxy * dmaxy + dminy * dminy;
                    if (sqd >= rsq) continue;
                    if (!heap.push(child, sqp)) return 0xffffffffu;
                } else {
                    if (child == tid || child == nb0 || child == nb1) continue;
                    const float2 q = A.P[child];
                    const float dx = p1.x - q.x, dy = p1.y - q.y;
                    const float d2 = dx * dx + dy * dy;
                    if (d2 < best && m2_incircle_sos(tid, nb0, nb1, child, A.P, A.stats) > 0.0f) {
                        g = child;
                        best = d2;
                    }
                }
            }
        }
        if (g != tid) break;
        st.set_elem(et, st.elem(nds));
        st.set_elem(nds, el);
        ++nds;
    }
    return g;
}

// Mesh::insert_by_circumsphere_checking (mesh.cpp:416-701): up to 32 rounds
// of (find one node inside a circumcircle, insert it) per node.
__global__ __launch_bounds__(kM2Block) void m2_iterative_kernel(M2Args A, const uint32_t* __restrict__ root,
                                                                const uint32_t* __restrict__ children,
                                                                const BoxD<float, 2>* __restrict__ data) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= A.n) return;
    uint32_t nloc = A.nlocal[tid];
    if (nloc & kFrozen) return;
    uint32_t T = A.nelem[tid];
    const M2Star st = m2_star(A.elem, A.neig, tid);
    const uint32_t r0 = *root;
    uint32_t nds = 0, r = 0;
    for (; r < kM2Rounds; ++r) {
        const uint32_t g = m2_find(A, st, tid, T, nds, r0, children, data);
        if (g == 0xffffffffu) {
            m2_freeze(A, tid, T, nloc);
            return;
        }
        if (g == tid) break;
        if (!m2_insert(A, st, tid, g, nds, T, nloc)) {
            m2_freeze(A, tid, T, nloc);
            return;
        }
    }
    atomicMax(&A.stats[kM2MaxRounds], r < kM2Rounds ? r + 1 : kM2Rounds);
    A.nelem[tid] = T;
    A.nlocal[tid] = nloc;
}

// Mesh::remove_super_nodes (mesh.cpp:704-745): drop the triangles on a super
// node; count the triangles this node outputs (it has the lowest index).
__global__ __launch_bounds__(kM2Block) void m2_remove_kernel(M2Args A, uint32_t* __restrict__ nelem_out) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= A.n) return;
    const M2Star st = m2_star(A.elem, A.neig, tid);
    uint32_t T = A.nelem[tid], ti = 0, nout = 0, nl = 0;
    while (ti < T) {
        const uint16_t el = st.elem(ti);
        const uint32_t ea = el & 255u, eb = el >> 8;
        if (ea < kM2Inf || eb < kM2Inf) {
            --T;
            st.set_elem(ti, st.elem(T));
            continue;
        }
        const uint32_t nb0 = st.glob(ea, A.n), nb1 = st.glob(eb, A.n);
        ++nl;
        nout += !((tid > nb0) || (tid > nb1));
        ++ti;
    }
    A.nelem[tid] = nl;
    nelem_out[tid] = nout;
}

// Mesh::compress_into_global (mesh.cpp:749-795).
__global__ __launch_bounds__(256) void m2_compress_kernel(M2Args A, const uint32_t* __restrict__ off,
                                                          uint32_t* __restrict__ out) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= A.n) return;
    const M2Star st = m2_star(A.elem, A.neig, tid);
    const uint32_t T = A.nelem[tid];
    uint32_t ti = 0;
    for (uint32_t i = off[tid]; i < off[tid + 1]; ++i) {
        uint32_t nb0 = 0, nb1 = 0;
        while (ti < T) {
            const uint16_t el = st.elem(ti++);
            nb0 = st.glob(el & 255u, A.n);
            nb1 = st.glob(el >> 8, A.n);
            if (!(nb0 < tid || nb1 < tid)) break;
        }
        out[3ull * i] = tid;
        out[3ull * i + 1] = nb0;
        out[3ull * i + 2] = nb1;
    }
}

}  // namespace

struct Mesh2DGpu {
    uint32_t n = 0;
    bool ready = false;   // init done
    bool cells = false;   // kd cells computed
    DBuf<Sph4<float>> obj;  // input nodes
    Agglo<float, 2> bvh;
    DBuf<float2> P;
    DBuf<uint16_t> elem;
    DBuf<uint32_t> neig, nelem, nlocal, nelem_out, off, out, stats;
    DBuf<BoxD<float, 2>> cell;
    DBuf<uint64_t> scan_ws;
    uint32_t n_elems = 0;
    float ms[5] = {0, 0, 0, 0, 0};  // init, morton, bvh, iterative, remove+compress

    M2Args args() {
        return {n, P.p, elem.p, neig.p, nelem.p, nlocal.p, stats.p};
    }
    void set_nodes(uint32_t nn, const float* xy) {
        if (nn < 2) throw std::invalid_argument("Mesh2D needs at least 2 nodes");
        std::vector<Sph4<float>> h(nn);
        for (uint32_t i = 0; i < nn; ++i) h[i] = {{xy[2 * i], xy[2 * i + 1], 0.0f}, 0.0f};
        obj.resize(nn);
        h2d(obj.p, h.data(), nn);
        n = nn;
        ready = false;
    }
    template <class F> void timed(int k, F&& f) {
        hipStream_t s = ctx().stream;
        hipEvent_t e0, e1;
        SA_HIP(hipEventCreate(&e0));
        SA_HIP(hipEventCreate(&e1));
        SA_HIP(hipEventRecord(e0, s));
        f(s);
        SA_HIP(hipEventRecord(e1, s));
        SA_HIP(hipEventSynchronize(e1));
        SA_HIP(hipEventElapsedTime(&ms[k], e0, e1));
        SA_HIP(hipEventDestroy(e0));
        SA_HIP(hipEventDestroy(e1));
        check();
    }
    void check() {
        const uint32_t f = read_scalar(ctx().d_flags + kFlagError);
        if (f) {
            fill<uint32_t>(ctx().d_flags, 1, 0u, ctx().stream);
            throw std::runtime_error("device error flag raised in Mesh2D");
        }
    }
    void need_init() const {
        if (!ready) throw std::logic_error("Mesh2D: set_nodes and init first");
    }
    void init() {
        if (!n) throw std::logic_error("Mesh2D: set_nodes first");
        timed(0, [&](hipStream_t s) {
            bvh.build(obj.p, n, 32, kLeafSphere, s);
            const size_t nblk = (n + kM2Block - 1) / kM2Block;
            P.resize(n + kM2Inf);
            elem.resize(nblk * kM2Cap * kM2Block);
            neig.resize(nblk * kM2Cap * kM2Block);
            nelem.resize(n);
            nlocal.resize(n);
            nelem_out.resize(n + 1);
            off.resize(n + 1);
            stats.resize(kM2StatWords);
            fill<uint32_t>(stats.p, kM2StatWords, 0u, s);
            hipLaunchKernelGGL(m2_init_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, bvh.base.obj_m.p,
                               bvh.base.gbox.p, P.p, elem.p, neig.p, nelem.p, nlocal.p, nelem_out.p);
        });
        ready = true;
        cells = false;
        n_elems = 0;
    }
    void insert_morton() {
        need_init();
        timed(1, [&](hipStream_t s) {
            hipLaunchKernelGGL(m2_morton_kernel, dim3(grid_for(n, kM2Block)), dim3(kM2Block), 0, s, args());
        });
    }
    void insert_bvh() {
        need_init();
        timed(2, [&](hipStream_t s) {
            cell.resize(2ull * n);
            hipLaunchKernelGGL(m2_cells_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, bvh.root.p,
                               bvh.children.p, bvh.split.p, bvh.range.p, bvh.base.codes_sorted, bvh.data.p,
                               bvh.base.gbox.p, cell.p, ctx().d_flags);
            hipLaunchKernelGGL(m2_bvh_kernel, dim3(grid_for(n, kM2Block)), dim3(kM2Block), 0, s, args(), bvh.root.p,
                               bvh.children.p, bvh.range.p, cell.p);
        });
        cells = true;
    }
    void insert_iterative() {
        need_init();
        timed(3, [&](hipStream_t s) {
            hipLaunchKernelGGL(m2_iterative_kernel, dim3(grid_for(n, kM2Block)), dim3(kM2Block), 0, s, args(),
                               bvh.root.p, bvh.children.p, bvh.data.p);
        });
    }
    void remove_super_nodes() {
        need_init();
        timed(4, [&](hipStream_t s) {
            hipLaunchKernelGGL(m2_remove_kernel, dim3(grid_for(n, kM2Block)), dim3(kM2Block), 0, s, args(),
                               nelem_out.p);
        });
    }
    // Side effect of get_nelem in the reference (mesh.cpp:849-852).
    uint32_t compress() {
        need_init();
        hipStream_t s = ctx().stream;
        scan_ws.resize(scan_workspace_words(n));
        exclusive_scan_u32(nelem_out.p, off.p, n, scan_ws.p, s);
        n_elems = read_scalar(off.p + n);
        out.resize(3ull * (n_elems ? n_elems : 1));
        hipLaunchKernelGGL(m2_compress_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, args(), off.p, out.p);
        SA_HIP(hipStreamSynchronize(s));
        return n_elems;
    }
};

}  // namespace stream_amd

using namespace stream_amd;

struct Mesh2D {
    Mesh2DGpu m;
    int32_t status = 0;
};

namespace {
template <class F> int32_t run_m2(Mesh2D* h, F&& f) {
    if (!h) {
        set_last_error("null Mesh2D handle");
        return kErrArgument;
    }
    h->status = guarded(f);
    return h->status;
}
}  // namespace

extern "C" {

// src/mesh/mesh.cpp:806-808
Mesh2D* Mesh2D_create(void) {
    try {
        (void)ctx();
        return new Mesh2D();
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return nullptr;
    }
}
// mesh.cpp:811-813
void Mesh2D_destroy(Mesh2D* h) { delete h; }
// mesh.cpp:816-823: copy the nodes (n x 2 floats) to the device.
void Mesh2D_set_nodes(Mesh2D* h, uint32_t nnodes, const float* nodes) {
    run_m2(h, [&] { h->m.set_nodes(nnodes, nodes); });
}
// mesh.cpp:826-828 (Mesh::init, :68-165)
void Mesh2D_init(Mesh2D* h) {
    run_m2(h, [&] { h->m.init(); });
}
// mesh.cpp:831-833 (:167-252)
void Mesh2D_insert_morton_neighbors(Mesh2D* h) {
    run_m2(h, [&] { h->m.insert_morton(); });
}
// mesh.cpp:837-839 (:254-413)
void Mesh2D_insert_BVH_neighbors(Mesh2D* h) {
    run_m2(h, [&] { h->m.insert_bvh(); });
}
// mesh.cpp:841-843 (:416-701)
void Mesh2D_remove_super_nodes(Mesh2D* h) {
    run_m2(h, [&] { h->m.remove_super_nodes(); });
}
// mesh.cpp:845-847
void Mesh2D_insert_iterative(Mesh2D* h) {
    run_m2(h, [&] { h->m.insert_iterative(); });
}
// mesh.cpp:849-852: compresses the stars into the global triangle list.
uint32_t Mesh2D_get_nelem(Mesh2D* h) {
    uint32_t k = 0;
    run_m2(h, [&] { k = h->m.compress(); });
    return k;
}
// mesh.cpp:854-858: triangles (n_elem x 3, Morton-order node ids).
void Mesh2D_get_elem(Mesh2D* h, uint32_t* elems) {
    run_m2(h, [&] { d2h(elems, h->m.out.p, 3ull * h->m.n_elems); });
}
// mesh.cpp:860-862: Morton-ordered nodes (n x 2).
void Mesh2D_get_ordered_nodes(Mesh2D* h, float* nodes) {
    run_m2(h, [&] {
        h->m.need_init();
        d2h(reinterpret_cast<float2*>(nodes), h->m.P.p, h->m.n);
    });
}
// Status of the last call on h (0 = ok; Stream_last_error has the text).
int32_t Mesh2D_last_status(const Mesh2D* h) { return h ? h->status : kErrArgument; }
// Counters: [0] nodes frozen on a capacity overflow, [1] most rounds any
// node ran in insert_iterative, [2] exact incircle evaluations; ms [5]: GPU
// time of init, insert_morton_neighbors, insert_BVH_neighbors,
// insert_iterative, remove_super_nodes (last call of each).
int32_t Mesh2D_get_stats(Mesh2D* h, uint32_t* counters, float* ms) {
    return run_m2(h, [&] {
        h->m.need_init();
        if (counters) d2h(counters, h->m.stats.p, 3);
        if (ms)
            for (int k = 0; k < 5; ++k) ms[k] = h->m.ms[k];
    });
}
// Stars for tests: elem [n][64] (a | b << 8, first nelem valid), nelem [n],
// neig [n][64] (local index j >= 3 at j - 3, first nlocal - 3 valid),
// nlocal [n] (bit 31 = frozen), permutation [n] (Morton slot -> input
// index).  Any pointer may be null.
int32_t Mesh2D_get_stars(Mesh2D* h, uint16_t* elem, uint32_t* nelem, uint32_t* neig, uint32_t* nlocal,
                         uint32_t* perm) {
    return run_m2(h, [&] {
        h->m.need_init();
        const uint32_t n = h->m.n;
        const size_t nblk = (n + kM2Block - 1) / kM2Block;
        if (elem || neig) {
            std::vector<uint16_t> e(nblk * kM2Cap * kM2Block);
            std::vector<uint32_t> g(nblk * kM2Cap * kM2Block);
            d2h(e.data(), h->m.elem.p, e.size());
            d2h(g.data(), h->m.neig.p, g.size());
            for (uint32_t t = 0; t < n; ++t) {
                const size_t base = (size_t)(t / kM2Block) * kM2Cap * kM2Block + (t % kM2Block);
                for (uint32_t j = 0; j < kM2Cap; ++j) {
                    if (elem) elem[(size_t)t * kM2Cap + j] = e[base + (size_t)j * kM2Block];
                    if (neig) neig[(size_t)t * kM2Cap + j] = g[base + (size_t)j * kM2Block];
                }
            }
        }
        if (nelem) d2h(nelem, h->m.nelem.p, n);
        if (nlocal) d2h(nlocal, h->m.nlocal.p, n);
        if (perm) d2h(perm, h->m.bvh.base.perm, n);
    });
}
// kd cells of the BVH nodes [2n-1] rows (lo0, lo1, hi0, hi1) after
// insert_BVH_neighbors.
int32_t Mesh2D_get_cells(Mesh2D* h, float* cells) {
    return run_m2(h, [&] {
        if (!h->m.cells) throw std::logic_error("Mesh2D: cells exist after insert_BVH_neighbors");
        d2h(reinterpret_cast<BoxD<float, 2>*>(cells), h->m.cell.p, 2ull * h->m.n - 1);
    });
}

}  // extern "C"
