// CPU oracle of the full 2D Delaunay mesher (Mesh2D).
//
// TEST INFRASTRUCTURE ONLY: compiled into oracle/lib/liboracle.so for tests/
// and the smoke check.  Restates src/mesh/mesh.cpp:45-796 and
// include/mesh/mesh.hpp:84-332 (fp_tt = float, the reference's only build).
//
// Every node keeps a local star: up to 32 triangles (node, neighbour a,
// neighbour b) over local neighbour indices, 0..2 being the three super-
// triangle vertices and 3.. the inserted neighbours.  The phases (Morton
// neighbours, BVH-cell neighbours, iterative circumcircle search, super-node
// removal, compression) all act on one node's star at a time and read only
// the coordinates and the agglomerative BVH, so running the nodes one after
// the other gives the reference's result for any scheduling.
//
// Capacity.  The reference keeps 32 triangle slots per node without a
// check, its cavity bit set is 32 bits wide (`& 31`, mesh.cpp:230-245) so
// local indices >= 32 alias, and its search heap holds 32 entries with no
// bound check (traversal_stack.hpp:47); past those limits its result is
// undefined (uniform clouds of 2e5 points already overflow the heap, 1e6
// points reach local index 32).  Here the triangle slots, local indices
// (64-bit cavity set) and heap entries go to 64; below the reference's
// limits every step is the same, so the outputs are the reference's.  A
// node that would pass 64 is frozen and counted in `overflow`.
#pragma once
#include "oracle.hpp"

namespace orc {

struct Mesh2D {
    static constexpr uint32_t kInf = 3;          // mesh.hpp:91
    static constexpr uint32_t kMaxElem = 64;     // reference: 32 (n_init_local_elem, mesh.hpp:112)
    static constexpr uint32_t kMaxLocal = 64;    // reference: 32 (cavity bit set, mesh.cpp:230)
    static constexpr uint32_t kRounds = 32;      // cur_max_nneig, mesh.cpp:456
    static constexpr uint32_t kMortonInsert = 10;  // n_init_insert, mesh.hpp:104
    static constexpr int kHeapCap = 64;          // reference: 32 (TraversalMinHeap, mesh.cpp:42)
    static constexpr int kStackCap = 32;         // TraversalStack<.., 32>, mesh.cpp:43

    struct Star {
        Edge8 e[kMaxElem];
        uint32_t neig[kMaxLocal];  // local index j >= 3 at j - 3 (mesh.hpp:233-247)
        uint32_t nelem = 0, nlocal = 0;
        bool frozen = false;
    };

    uint32_t n = 0;
    LbvhA<float, 2> bvh;
    std::vector<float> xy;  // Morton-ordered nodes + 3 super nodes, 2 per node
    std::vector<Star> star;
    std::vector<Box<float, 2>> cell;  // implicit kd cells of the BVH nodes
    std::vector<uint32_t> nelem_out;  // [n + 1]
    std::vector<uint32_t> elem;       // compressed triangles, 3 per
    uint32_t n_elems = 0;
    uint64_t overflow = 0;
    uint32_t last_rounds = 0;

    const float* pt(uint32_t g) const { return &xy[2ull * g]; }
    uint32_t global(const Star& s, uint32_t j) const { return j >= kInf ? s.neig[j - kInf] : n + j; }

    // Mesh::init (mesh.cpp:68-165): agglomerative BVH over the points,
    // super-triangle at avg +- (max extent + 10), star = 3 triangles.
    void init(const float* pts, uint32_t npts) {
        n = npts;
        std::vector<Sphere<float, 2>> obj(n);
        for (uint32_t i = 0; i < n; ++i) obj[i] = {{pts[2 * i], pts[2 * i + 1]}, 0.0f};
        build_lbvha(bvh, obj, 32, [](const Sphere<float, 2>& s) { return sphere_box(s); });
        xy.assign(2ull * (n + kInf), 0.0f);
        for (uint32_t i = 0; i < n; ++i) {
            xy[2 * i] = bvh.obj_m[i].c[0];
            xy[2 * i + 1] = bvh.obj_m[i].c[1];
        }
        const Box<float, 2>& g = bvh.global;
        const float avg0 = 0.5f * (g.hi[0] + g.lo[0]), avg1 = 0.5f * (g.hi[1] + g.lo[1]);
        const float mr = std::fmax(g.hi[0] - g.lo[0], g.hi[1] - g.lo[1]) + 10.0f;
        const float inf[6] = {avg0 - mr, avg1 - mr, avg0 + mr, avg1 - mr, avg0, avg1 + mr};
        for (int k = 0; k < 6; ++k) xy[2ull * n + k] = inf[k];
        star.assign(n, Star());
        for (auto& s : star) {
            s.e[0] = {1, 2};
            s.e[1] = {2, 0};
            s.e[2] = {0, 1};
            s.nelem = 3;
            s.nlocal = kInf;
        }
        nelem_out.assign(n + 1, 0);
        elem.clear();
        n_elems = 0;
        overflow = 0;
    }

    // incircle_SoS (mesh.hpp:287-329).
    float incircle_sos(uint32_t i1, uint32_t i2, uint32_t i3, uint32_t i4) const {
        const float *pa = pt(i1), *pb = pt(i2), *pc = pt(i3), *pd = pt(i4);
        const float orientation = orient2d(pa, pb, pc);
        float err;
        float det = incircle_filter(pa, pb, pc, pd, &err);
        if (!((det > err) || (-det > err))) {
            bool swap = false;
            auto cs = [&](uint32_t& x, uint32_t& y, const float*& px, const float*& py) {
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
            det = (float)exact::incircle_sign(pa, pb, pc, pd);
            int depth = 0;
            while (det == 0.0f) {
                ++depth;
                if (depth == 1)
                    det = orient2d(pb, pc, pd);
                else if (depth == 2)
                    det = -orient2d(pa, pc, pd);
                else
                    det = ((int)swap + (int)(orientation < 0.0f)) == 1 ? 1.0f : -1.0f;
            }
            if (swap) det = -det;
        }
        if (orientation < 0.0f) det = -det;
        return det;
    }

    // Bowyer-Watson insertion of global node g into the star of tid, over
    // triangles [first, nelem) (mesh.cpp:210-245, 359-395, 639-675).
    // Returns 1 when the star overflows (it is then frozen), else 0.
    uint32_t insert(uint32_t tid, uint32_t g, uint32_t first) {
        Star& s = star[tid];
        if (s.frozen) return 0;
        const uint32_t loc = s.nlocal;  // <= kMaxLocal: it only grows below it
        s.neig[loc - kInf] = g;
        uint64_t shell = 0;
        uint32_t ti = first;
        // the cavity is cut on a copy so that a star that would overflow is
        // left as it was
        Edge8 e[kMaxElem];
        std::memcpy(e, s.e, sizeof(e));
        uint32_t T = s.nelem;
        while (ti < T) {
            const Edge8 el = e[ti];
            if (incircle_sos(tid, global(s, el.a), global(s, el.b), g) > 0.0f) {
                shell ^= 1ull << el.a;
                shell ^= 1ull << el.b;
                --T;
                e[ti] = e[T];
            } else {
                ++ti;
            }
        }
        if (shell) {
            if (loc >= kMaxLocal || T + 2 > kMaxElem) {
                s.frozen = true;
                return 1;
            }
            e[T++] = {(uint8_t)__builtin_ctzll(shell), (uint8_t)loc};
            e[T++] = {(uint8_t)(63 - __builtin_clzll(shell)), (uint8_t)loc};
            s.nlocal = loc + 1;
        }
        std::memcpy(s.e, e, sizeof(e));
        s.nelem = T;
        return 0;
    }

    // Mesh::insert_morton_neighbors (mesh.cpp:167-252).
    void insert_morton_neighbors() {
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : overflow)
        for (uint32_t tid = 0; tid < n; ++tid) {
            const uint32_t h = kMortonInsert / 2;
            uint32_t lo = tid - h, hi = tid + h;
            if (tid < h) {
                lo = 0;
                hi = n < kMortonInsert ? n : kMortonInsert + 1;
            }
            if (tid + h >= n) {
                lo = n < kMortonInsert ? 0 : n - kMortonInsert;
                hi = n;
            }
            for (uint32_t g = lo; g < hi; ++g)
                if (g != tid) overflow += insert(tid, g, 0);
        }
    }

    // Cells of the implicit kd-tree the BVH splits describe
    // (mesh.cpp:274-311): descend from the root towards leaf tid, cutting
    // the global box at the mid-plane between the two children's boxes on
    // the axis of the highest differing Morton bit.
    void compute_cells() {
        cell.assign(2ull * n - 1, Box<float, 2>::empty());
        if (n < 2) return;
        cell[bvh.root] = bvh.global;  // never read; the reference leaves it unset
#pragma omp parallel for schedule(static)
        for (uint32_t tid = 0; tid < n; ++tid) {
            Box<float, 2> bb = bvh.global;
            uint32_t cur = bvh.root;
            while (true) {
                const uint32_t k = cur - n;
                const uint32_t sp = bvh.split[k];
                const int bit = clz64(bvh.codes_sorted[sp] ^ bvh.codes_sorted[sp + 1]);
                const int axis = (bit + 1) % 2;
                const uint32_t c0 = bvh.children[2 * k], c1 = bvh.children[2 * k + 1];
                const float coord = 0.5f * (bvh.data[c0].hi[axis] + bvh.data[c1].lo[axis]);
                const uint32_t child = tid > sp ? c1 : c0;
                if (child == c0)
                    bb.hi[axis] = coord;
                else
                    bb.lo[axis] = coord;
                if (child < n || bvh.range[2 * (child - n)] == tid) cell[child] = bb;
                if (child < n) break;
                cur = child;
            }
        }
    }

    // Mesh::insert_BVH_neighbors (mesh.cpp:313-413): every leaf whose cell
    // touches the node's cell, in stack order; an internal node met with a
    // full stack has its whole leaf range inserted.
    void insert_bvh_neighbors() {
        if (n < 2) return;
        compute_cells();
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : overflow)
        for (uint32_t tid = 0; tid < n; ++tid) {
            const Box<float, 2> bl = cell[tid];
            uint32_t stack[kStackCap];
            int len = 0;
            stack[len++] = bvh.root;
            while (len) {
                const uint32_t cur = stack[--len];
                for (int ch = 0; ch < 2; ++ch) {
                    const uint32_t child = bvh.children[2 * (cur - n) + ch];
                    const Box<float, 2>& bn = cell[child];
                    bool nb = true;
                    nb &= !(bl.hi[0] < bn.lo[0] || bl.lo[0] > bn.hi[0]);
                    nb &= !(bl.hi[1] < bn.lo[1] || bl.lo[1] > bn.hi[1]);
                    if (!nb) continue;
                    const bool leaf = child < n || len >= kStackCap;
                    if (!leaf) {
                        stack[len++] = child;
                    } else {
                        const uint32_t rmin = child < n ? child : bvh.range[2 * (child - n)];
                        const uint32_t rmax = child < n ? child : bvh.range[2 * (child - n) + 1];
                        for (uint32_t o = rmin; o <= rmax; ++o)
                            if (o != tid) overflow += insert(tid, o, 0);
                    }
                }
            }
        }
    }

    // Min-heap with the reference's sift rules (traversal_stack.hpp:31-90).
    struct TraversalHeapO {
        struct Pair {
            uint32_t first;
            float second;
        };
        int len = 0;
        Pair item[kHeapCap];
        bool push(uint32_t idx, float cost) {
            if (len >= kHeapCap) return false;
            item[len++] = {idx, cost};
            int i = len - 1, p = (i - 1) >> 1;
            while (i > 0 && item[p].second > item[i].second) {
                std::swap(item[i], item[p]);
                i = p;
                p = (i - 1) >> 1;
            }
            return true;
        }
        void pop() {
            item[0] = item[--len];
            int i = 0;
            while (true) {
                const int l = 2 * i + 1, r = 2 * i + 2;
                int s = i;
                if (l < len && item[l].second < item[s].second) s = l;
                if (r < len && item[r].second < item[s].second) s = r;
                if (s == i) break;
                std::swap(item[i], item[s]);
                i = s;
            }
        }
    };

    // Nearest node to tid strictly inside the circumcircle of the star's
    // triangles from `nds` on (mesh.cpp:465-566).  Triangles found empty
    // move to the front.  The heap persists across the triangles of one
    // call, as in the reference.  Returns tid when every triangle is empty;
    // sets `bad` on heap overflow.
    uint32_t find(uint32_t tid, uint32_t& nds, bool& bad) {
        Star& s = star[tid];
        TraversalHeapO heap;
        uint32_t g = tid;
        const float* p1 = pt(tid);
        for (uint32_t et = nds; et < s.nelem; ++et) {
            const Edge8 el = s.e[et];
            const uint32_t nb0 = global(s, el.a), nb1 = global(s, el.b);
            g = tid;
            const float* p2 = pt(nb0);
            const float* p3 = pt(nb1);
            const float bax = p2[0] - p1[0], bay = p2[1] - p1[1];
            const float cax = p3[0] - p1[0], cay = p3[1] - p1[1];
            const float cbx = p3[0] - p2[0], cby = p3[1] - p2[1];
            const float l1 = bax * bax + bay * bay;
            const float l2 = cax * cax + cay * cay;
            const float l3 = cbx * cbx + cby * cby;
            const float det = bax * cay - bay * cax;
            const float rsq = 1.0000001f * ((l1 * l2 * l3) / (4.f * det * det));
            const float ox = p3[0] - 0.5f * (l2 * cby - l3 * cay) / det;
            const float oy = p3[1] - 0.5f * (l3 * cax - l2 * cbx) / det;
            if (!heap.push(bvh.root, 0.0f)) {
                bad = true;
                return tid;
            }
            float best = 4.f * rsq;
            while (heap.len) {
                const auto top = heap.item[0];
                if (top.second >= best) break;
                heap.pop();
                for (int ch = 0; ch < 2; ++ch) {
                    const uint32_t child = bvh.children[2 * (top.first - n) + ch];
                    if (child >= n) {
                        const Box<float, 2>& b = bvh.data[child];
                        float dmaxx = std::fmax(p1[0] - b.hi[0], 0.0f);
                        float dmaxy = std::fmax(p1[1] - b.hi[1], 0.0f);
                        float dminx = std::fmax(b.lo[0] - p1[0], 0.0f);
                        float dminy = std::fmax(b.lo[1] - p1[1], 0.0f);
                        const float sqp = dmaxx * dmaxx + dminx * dminx + dmaxy * dmaxy + dminy * dminy;
                        if (sqp >= best) continue;
                        dmaxx = std::fmax(ox - b.hi[0], 0.0f);
                        dmaxy = std::fmax(oy - b.hi[1], 0.0f);
                        dminx = std::fmax(b.lo[0] - ox, 0.0f);
                        dminy = std::fmax(b.lo[1] - oy, 0.0f);
                        const float sqd = dmaxx * dmaxx + dminx * dminx + dmaxy * dmaxy + dminy * dminy;
                        if (sqd >= rsq) continue;
                        if (!heap.push(child, sqp)) {
                            bad = true;
                            return tid;
                        }
                    } else {
                        if (child == tid || child == nb0 || child == nb1) continue;
                        const float* q = pt(child);
                        const float dx = p1[0] - q[0], dy = p1[1] - q[1];
                        const float d2 = dx * dx + dy * dy;
                        if (d2 < best && incircle_sos(tid, nb0, nb1, child) > 0.0f) {
                            g = child;
                            best = d2;
                        }
                    }
                }
            }
            if (g != tid) break;
            s.e[et] = s.e[nds];
            s.e[nds] = el;
            ++nds;
        }
        return g;
    }

    // Mesh::insert_by_circumsphere_checking (mesh.cpp:416-701): at most 32
    // rounds of (find one node inside a circumcircle, insert it); a node
    // whose triangles are all empty is done.  Rounds are independent per
    // node, so each node runs its own loop here.
    void insert_iterative() {
        uint32_t max_rounds = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : overflow) reduction(max : max_rounds)
        for (uint32_t tid = 0; tid < n; ++tid) {
            uint32_t nds = 0, r = 0;
            for (; r < kRounds; ++r) {
                Star& s = star[tid];
                if (s.frozen) break;
                bool bad = false;
                const uint32_t g = find(tid, nds, bad);
                if (bad) {
                    s.frozen = true;
                    ++overflow;
                    break;
                }
                if (g == tid) break;
                overflow += insert(tid, g, nds);
            }
            max_rounds = std::max(max_rounds, r + (r < kRounds ? 1u : 0u));
        }
        last_rounds = max_rounds;
    }

    // Mesh::remove_super_nodes (mesh.cpp:704-745).
    void remove_super_nodes() {
        nelem_out.assign(n + 1, 0);
#pragma omp parallel for schedule(static)
        for (uint32_t tid = 0; tid < n; ++tid) {
            Star& s = star[tid];
            uint32_t T = s.nelem, ti = 0, nout = 0, nloc = 0;
            while (ti < T) {
                const Edge8 el = s.e[ti];
                if (el.a < kInf || el.b < kInf) {
                    --T;
                    s.e[ti] = s.e[T];
                    continue;
                }
                const uint32_t nb0 = global(s, el.a), nb1 = global(s, el.b);
                ++nloc;
                nout += !((tid > nb0) || (tid > nb1));
                ++ti;
            }
            s.nelem = nloc;
            nelem_out[tid + 1] = nout;
        }
    }

    // Mesh::compress_into_global (mesh.cpp:749-795): each triangle once, from
    // its lowest-index node, in star order.
    void compress() {
        std::vector<uint32_t> off(n + 1, 0);
        for (uint32_t i = 0; i < n; ++i) off[i + 1] = off[i] + nelem_out[i + 1];
        n_elems = off[n];
        elem.assign(3ull * n_elems, 0);
        for (uint32_t tid = 0; tid < n; ++tid) {
            const Star& s = star[tid];
            uint32_t ti = 0;
            for (uint32_t i = off[tid]; i < off[tid + 1]; ++i) {
                uint32_t nb0, nb1;
                do {
                    const Edge8 el = s.e[ti];
                    nb0 = global(s, el.a);
                    nb1 = global(s, el.b);
                    ++ti;
                } while (nb0 < tid || nb1 < tid);
                elem[3ull * i] = tid;
                elem[3ull * i + 1] = nb0;
                elem[3ull * i + 2] = nb1;
            }
        }
    }
};

}  // namespace orc
