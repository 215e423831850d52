// C entry points of the CPU oracle, for ctypes.  TEST INFRASTRUCTURE ONLY:
// loaded by tests/, the smoke check and bench.py's cpu_baseline leg.
#include <chrono>
#include <cstdio>

#include "oracle.hpp"
#include "oracle3d.hpp"
#include "mesh2d.hpp"

using namespace orc;

namespace {

template <class S, int D>
std::vector<Sphere<S, D>> pack(uint32_t n, const S* xyr) {
    std::vector<Sphere<S, D>> v(n);
    for (uint32_t i = 0; i < n; ++i) {
        for (int a = 0; a < D; ++a) v[i].c[a] = xyr[(D + 1) * i + a];
        v[i].r = xyr[(D + 1) * i + D];
    }
    return v;
}

// LBVH over spheres; leaf_mode 0 = centre box (alpha-shape functor),
// 1 = sphere box (ComputeBB_Functor).
template <class S, int D>
int lbvh_build(uint32_t n, const S* xyr, int bits, int leaf_mode, uint64_t* codes, uint64_t* codes_sorted,
               uint32_t* perm, int32_t* sep, uint8_t* cl, uint32_t* leaf_parent, uint32_t* internal_parent,
               S* internal_box, S* global_box) {
    if (n < 2) return -1;
    auto obj = pack<S, D>(n, xyr);
    Lbvh<S, D> t;
    if (leaf_mode == 0)
        build_lbvh(t, obj, bits, [](const Sphere<S, D>& s) { return centre_box(s); });
    else
        build_lbvh(t, obj, bits, [](const Sphere<S, D>& s) { return sphere_box(s); });
    for (uint32_t i = 0; i < n; ++i) {
        codes[i] = t.codes[i];
        codes_sorted[i] = t.codes_sorted[i];
        perm[i] = t.perm[i];
        leaf_parent[i] = t.leaf_parent[i];
    }
    for (uint32_t i = 0; i + 1 < n; ++i) {
        sep[i] = t.sep[i];
        cl[i] = t.child_is_leaf[i];
        internal_parent[i] = t.internal_parent[i];
        for (int a = 0; a < D; ++a) {
            internal_box[2 * D * i + a] = t.internal_box[i].lo[a];
            internal_box[2 * D * i + D + a] = t.internal_box[i].hi[a];
        }
    }
    for (int a = 0; a < D; ++a) {
        global_box[a] = t.global.lo[a];
        global_box[D + a] = t.global.hi[a];
    }
    return 0;
}

// Sphere-sphere collision count over a sphere-box LBVH, the query of
// tests/geometry/lbvh_simple_collision.cpp:100-140 (per query, excluding self).
template <class S, int D>
uint64_t lbvh_collisions(uint32_t n, const S* xyr, int bits) {
    auto obj = pack<S, D>(n, xyr);
    Lbvh<S, D> t;
    build_lbvh(t, obj, bits, [](const Sphere<S, D>& s) { return sphere_box(s); });
    uint64_t total = 0;
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 256)
    for (uint32_t tid = 0; tid < n; ++tid) {
        int stack[64];
        int ss = 0;
        stack[ss++] = 0;
        const Sphere<S, D> q = t.obj_m[tid];
        uint64_t c = 0;
        while (ss) {
            const int cur = stack[--ss];
            const int sepv = t.sep[cur];
            const uint8_t cl = t.child_is_leaf[cur];
            for (int ch = 0; ch < 2; ++ch) {
                if (!(cl & (ch + 1))) {
                    const Box<S, D>& b = t.internal_box[sepv + ch];
                    bool in = true;
                    for (int a = 0; a < D; ++a) {
                        in &= q.c[a] - q.r < b.hi[a];
                        in &= q.c[a] + q.r > b.lo[a];
                    }
                    if (in) stack[ss++] = sepv + ch;
                } else {
                    const Sphere<S, D>& o = t.obj_m[sepv + ch];
                    const S rs = q.r + o.r;
                    S d2 = 0;
                    for (int a = 0; a < D; ++a) {
                        const S dd = q.c[a] - o.c[a];
                        d2 += dd * dd;
                    }
                    if (d2 <= rs * rs && (uint32_t)(sepv + ch) != tid) ++c;
                }
            }
        }
        total += c;
    }
    return total;
}

template <class S>
void* ashape2d_run(uint32_t n, const S* xyr, int bits, int contract = 0) {
    auto* a = new AlphaShape2D<S>();
    a->morton_bits = bits;
    a->contract = contract;
    a->init(pack<S, 2>(n, xyr));
    a->compute();
    return a;
}

template <class S>
void get2d(AlphaShape2D<S>* a, uint32_t* elem, uint32_t* row, uint32_t* neig, uint8_t* edge_bnd,
                  uint8_t* node_bnd, uint32_t* perm, S* xy_m, uint32_t* nineig) {
    const uint32_t n = a->n;
    if (elem) std::memcpy(elem, a->elem.data(), sizeof(uint32_t) * 3ull * a->n_elems);
    if (row) std::memcpy(row, a->row.data(), sizeof(uint32_t) * (n + 1));
    if (neig) std::memcpy(neig, a->neig.data(), sizeof(uint32_t) * a->n_edges);
    if (edge_bnd) std::memcpy(edge_bnd, a->edge_is_bnd.data(), a->n_edges);
    if (node_bnd) std::memcpy(node_bnd, a->node_is_bnd.data(), n);
    if (perm) std::memcpy(perm, a->lbvh.perm.data(), sizeof(uint32_t) * n);
    if (xy_m)
        for (uint32_t i = 0; i < n; ++i) {
            xy_m[2 * i] = a->coords_m[i].c[0];
            xy_m[2 * i + 1] = a->coords_m[i].c[1];
        }
    if (nineig)
        for (uint32_t i = 0; i < n; ++i) nineig[i] = a->nineig[i];
}
template <class S>
void* ashape3d_run(uint32_t n, const S* xyzr, int bits, int contract = 0) {
    auto* a = new AlphaShape3D<S>();
    a->morton_bits = bits;
    a->contract = contract;
    a->init(pack<S, 3>(n, xyzr));
    a->compute();
    return a;
}

template <class S>
void get3d(AlphaShape3D<S>* a, uint32_t* elem, uint32_t* row, uint32_t* neig, uint8_t* edge_bnd,
           uint8_t* node_bnd, uint32_t* perm, S* xyz_m, uint32_t* nineig) {
    const uint32_t n = a->n;
    if (elem) std::memcpy(elem, a->elem.data(), sizeof(uint32_t) * 4ull * a->n_elems);
    if (row) std::memcpy(row, a->row.data(), sizeof(uint32_t) * (n + 1));
    if (neig) std::memcpy(neig, a->neig.data(), sizeof(uint32_t) * a->n_edges);
    if (edge_bnd) std::memcpy(edge_bnd, a->edge_is_bnd.data(), a->n_edges);
    if (node_bnd) std::memcpy(node_bnd, a->node_is_bnd.data(), n);
    if (perm) std::memcpy(perm, a->lbvh.perm.data(), sizeof(uint32_t) * n);
    if (xyz_m)
        for (uint32_t i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) xyz_m[3 * i + c] = a->coords_m[i].c[c];
    if (nineig)
        for (uint32_t i = 0; i < n; ++i) nineig[i] = a->nineig[i];
}
template <int D>
int trbvh_build(uint32_t n, const float* xyr, int rounds, uint32_t* perm, uint32_t* children, uint32_t* range,
                uint32_t* parent, uint32_t* root, float* boxes, float* sah, uint32_t* subtree) {
    if (n < 2 || rounds < 0) return -1;
    Trbvh<D> T;
    build_trbvh<D>(T, pack<float, D>(n, xyr), rounds);
    if (perm) std::memcpy(perm, T.t.perm.data(), 4ull * n);
    if (children) std::memcpy(children, T.t.children.data(), 8ull * (n - 1));
    if (range) std::memcpy(range, T.t.range.data(), 8ull * (n - 1));
    if (parent) std::memcpy(parent, T.t.parent.data(), 4ull * (2 * n - 1));
    if (root) *root = T.t.root;
    if (boxes)
        for (uint32_t i = 0; i < 2 * n - 1; ++i)
            for (int a = 0; a < D; ++a) {
                boxes[2 * D * i + a] = T.t.data[i].lo[a];
                boxes[2 * D * i + D + a] = T.t.data[i].hi[a];
            }
    if (sah) std::memcpy(sah, T.sah.data(), 4ull * (2 * n - 1));
    if (subtree) std::memcpy(subtree, T.subtree.data(), 4ull * (2 * n - 1));
    return 0;
}
}  // namespace

extern "C" {

// ---- agglomerative LBVH: outputs as LBVHa2D_build_* ----
int orc_lbvha2d_build_f64(uint32_t n, const double* xyr, int bits, int leaf_mode, uint32_t* perm, uint32_t* children,
                          uint32_t* range, uint32_t* parent, uint32_t* root, double* boxes, uint32_t* split) {
    if (n < 2) return -1;
    auto obj = pack<double, 2>(n, xyr);
    LbvhA<double, 2> t;
    if (leaf_mode == 0)
        build_lbvha(t, obj, bits, [](const Sphere<double, 2>& s) { return centre_box(s); });
    else
        build_lbvha(t, obj, bits, [](const Sphere<double, 2>& s) { return sphere_box(s); });
    std::memcpy(perm, t.perm.data(), 4ull * n);
    std::memcpy(children, t.children.data(), 8ull * (n - 1));
    std::memcpy(range, t.range.data(), 8ull * (n - 1));
    std::memcpy(parent, t.parent.data(), 4ull * (2 * n - 1));
    *root = t.root;
    for (uint32_t i = 0; i < 2 * n - 1; ++i)
        for (int a = 0; a < 2; ++a) {
            boxes[4 * i + a] = t.data[i].lo[a];
            boxes[4 * i + 2 + a] = t.data[i].hi[a];
        }
    std::memcpy(split, t.split.data(), 4ull * (n - 1));
    return 0;
}
uint64_t orc_lbvha2d_collisions_f64(uint32_t n, const double* xyr, int bits, uint32_t leaf_size) {
    auto obj = pack<double, 2>(n, xyr);
    LbvhA<double, 2> t;
    build_lbvha(t, obj, bits, [](const Sphere<double, 2>& s) { return sphere_box(s); });
    return lbvha_collisions(t, leaf_size);
}
uint64_t orc_lbvha2d_collisions_f32(uint32_t n, const float* xyr, int bits, uint32_t leaf_size) {
    auto obj = pack<float, 2>(n, xyr);
    LbvhA<float, 2> t;
    build_lbvha(t, obj, bits, [](const Sphere<float, 2>& s) { return sphere_box(s); });
    return lbvha_collisions(t, leaf_size);
}
uint64_t orc_lbvha3d_collisions_f64(uint32_t n, const double* xyzr, int bits, uint32_t leaf_size) {
    auto obj = pack<double, 3>(n, xyzr);
    LbvhA<double, 3> t;
    build_lbvha(t, obj, bits, [](const Sphere<double, 3>& s) { return sphere_box(s); });
    return lbvha_collisions(t, leaf_size);
}

// ---- TRBVH (float, the reference's only build) ----
int orc_trbvh2d_build_f32(uint32_t n, const float* xyr, int rounds, uint32_t* perm, uint32_t* children,
                          uint32_t* range, uint32_t* parent, uint32_t* root, float* boxes, float* sah,
                          uint32_t* subtree) {
    return trbvh_build<2>(n, xyr, rounds, perm, children, range, parent, root, boxes, sah, subtree);
}
int orc_trbvh3d_build_f32(uint32_t n, const float* xyzr, int rounds, uint32_t* perm, uint32_t* children,
                          uint32_t* range, uint32_t* parent, uint32_t* root, float* boxes, float* sah,
                          uint32_t* subtree) {
    return trbvh_build<3>(n, xyzr, rounds, perm, children, range, parent, root, boxes, sah, subtree);
}
uint64_t orc_trbvh2d_collisions_f32(uint32_t n, const float* xyr, int rounds) {
    Trbvh<2> T;
    build_trbvh<2>(T, pack<float, 2>(n, xyr), rounds);
    return trbvh_collisions<2>(T);
}
uint64_t orc_trbvh3d_collisions_f32(uint32_t n, const float* xyzr, int rounds) {
    Trbvh<3> T;
    build_trbvh<3>(T, pack<float, 3>(n, xyzr), rounds);
    return trbvh_collisions<3>(T);
}
// The treelet optimiser's dynamic programme on 5 leaves given as box rows
// (lo..., hi...), leaf costs and subtree sizes (trbvh.hpp:460-533):
// copt/popt/size get 32 entries indexed by leaf-subset bit mask.
void orc_treelet_dp2d(const float* boxes, const float* cost, const uint32_t* lsize, float* copt, uint32_t* popt,
                      uint32_t* size) {
    Box<float, 2> b[kTreeletLeaves];
    for (int l = 0; l < kTreeletLeaves; ++l)
        for (int a = 0; a < 2; ++a) {
            b[l].lo[a] = boxes[4 * l + a];
            b[l].hi[a] = boxes[4 * l + 2 + a];
        }
    treelet_dp<2>(b, cost, lsize, copt, popt, size);
}

// ---- Mesh2D (full 2D Delaunay, float) ----
void* orc_mesh2d_create(uint32_t n, const float* xy) {
    auto* m = new Mesh2D();
    m->init(xy, n);
    return m;
}
void orc_mesh2d_free(void* h) { delete static_cast<Mesh2D*>(h); }
void orc_mesh2d_insert_morton(void* h) { static_cast<Mesh2D*>(h)->insert_morton_neighbors(); }
void orc_mesh2d_insert_bvh(void* h) { static_cast<Mesh2D*>(h)->insert_bvh_neighbors(); }
uint32_t orc_mesh2d_insert_iterative(void* h) {
    auto* m = static_cast<Mesh2D*>(h);
    m->insert_iterative();
    return m->last_rounds;
}
void orc_mesh2d_remove_super(void* h) { static_cast<Mesh2D*>(h)->remove_super_nodes(); }
uint32_t orc_mesh2d_nelem(void* h) {
    auto* m = static_cast<Mesh2D*>(h);
    m->compress();
    return m->n_elems;
}
void orc_mesh2d_elem(void* h, uint32_t* out) {
    auto* m = static_cast<Mesh2D*>(h);
    std::memcpy(out, m->elem.data(), 12ull * m->n_elems);
}
uint64_t orc_mesh2d_overflow(void* h) { return static_cast<Mesh2D*>(h)->overflow; }
// Stars: elem [n][64] as (a | b << 8), nelem [n], neig [n][64] (local index
// j >= 3 at j - 3), nlocal [n], ordered nodes [n][2] (any may be null).
void orc_mesh2d_stars(void* h, uint16_t* elem, uint32_t* nelem, uint32_t* neig, uint32_t* nlocal, float* xy) {
    auto* m = static_cast<Mesh2D*>(h);
    for (uint32_t i = 0; i < m->n; ++i) {
        const auto& st = m->star[i];
        if (elem)
            for (uint32_t k = 0; k < Mesh2D::kMaxElem; ++k)
                elem[64ull * i + k] = k < st.nelem ? (uint16_t)(st.e[k].a | (st.e[k].b << 8)) : 0;
        if (nelem) nelem[i] = st.nelem;
        if (neig)
            for (uint32_t k = 0; k < Mesh2D::kMaxLocal; ++k) neig[64ull * i + k] = k + 3 < st.nlocal ? st.neig[k] : 0;
        if (nlocal) nlocal[i] = st.nlocal;
        if (xy) {
            xy[2 * i] = m->xy[2 * i];
            xy[2 * i + 1] = m->xy[2 * i + 1];
        }
    }
}
// kd cells of the BVH nodes [2n-1] rows (lo0, lo1, hi0, hi1), computed on demand.
void orc_mesh2d_cells(void* h, float* out) {
    auto* m = static_cast<Mesh2D*>(h);
    m->compute_cells();
    for (size_t i = 0; i < m->cell.size(); ++i) {
        out[4 * i] = m->cell[i].lo[0];
        out[4 * i + 1] = m->cell[i].lo[1];
        out[4 * i + 2] = m->cell[i].hi[0];
        out[4 * i + 3] = m->cell[i].hi[1];
    }
}

// ---- 3D predicates ----
float orc_orient3d_f32(const float* a, const float* b, const float* c, const float* d) { return orient3d(a, b, c, d); }
double orc_orient3d_f64(const double* a, const double* b, const double* c, const double* d) {
    return orient3d(a, b, c, d);
}
int orc_orient3d_exact_f32(const float* a, const float* b, const float* c, const float* d) {
    return exact::orient3d_sign(a, b, c, d);
}
int orc_orient3d_exact_f64(const double* a, const double* b, const double* c, const double* d) {
    return exact::orient3d_sign(a, b, c, d);
}
float orc_insphere_filter_f32(const float* a, const float* b, const float* c, const float* d, const float* e,
                              float* err) {
    return insphere_filter(a, b, c, d, e, err);
}
double orc_insphere_filter_f64(const double* a, const double* b, const double* c, const double* d, const double* e,
                               double* err) {
    return insphere_filter(a, b, c, d, e, err);
}
int orc_insphere_exact_f32(const float* a, const float* b, const float* c, const float* d, const float* e) {
    return exact::insphere_sign(a, b, c, d, e);
}
int orc_insphere_exact_f64(const double* a, const double* b, const double* c, const double* d, const double* e) {
    return exact::insphere_sign(a, b, c, d, e);
}

// ---- P1 on tetrahedra: returns nnz; row/col/val/b must hold n+1 / cap / cap / n ----
int64_t orc_assemble_p1_tet(uint32_t n, const double* xyz, uint32_t ne, const uint32_t* elem, const uint8_t* bnd,
                            const double* f, uint64_t cap, uint32_t* row, uint32_t* col, double* val, double* b) {
    CsrD A;
    std::vector<double> bb;
    assemble_p1_tet(n, xyz, ne, elem, bnd, f, A, bb);
    if (A.col.size() > cap) return -(int64_t)A.col.size();
    std::memcpy(row, A.row.data(), sizeof(uint32_t) * (n + 1));
    std::memcpy(col, A.col.data(), sizeof(uint32_t) * A.col.size());
    std::memcpy(val, A.val.data(), sizeof(double) * A.val.size());
    std::memcpy(b, bb.data(), sizeof(double) * n);
    return (int64_t)A.col.size();
}

// ---- alpha-shape 3D ----
void* orc_ashape3d_run_f32(uint32_t n, const float* xyzr, int bits) { return ashape3d_run<float>(n, xyzr, bits); }
void* orc_ashape3d_run_f64(uint32_t n, const double* xyzr, int bits) { return ashape3d_run<double>(n, xyzr, bits); }
// the same with the inexact tests evaluated in fmac mode `contract`
void* orc_ashape3d_run_contract_f32(uint32_t n, const float* xyzr, int bits, int contract) {
    return ashape3d_run<float>(n, xyzr, bits, contract);
}
// the README-style FP32 run with a cavity-walk cap of `cap` steps, keeping the
// points of every exact insphere call (readable with orc_ashape3d_exact_points_f32)
void* orc_ashape3d_run_probe_f32(uint32_t n, const float* xyzr, int bits, int cap, int record) {
    auto* a = new AlphaShape3D<float>();
    a->morton_bits = bits;
    a->max_cavity_iter = cap;
    a->record_exact = record != 0;
    a->init(pack<float, 3>(n, xyzr));
    a->compute();
    return a;
}
uint64_t orc_ashape3d_exact_points_f32(void* h, float* out) {
    auto* a = static_cast<AlphaShape3D<float>*>(h);
    if (out) std::memcpy(out, a->exact_points.data(), sizeof(float) * a->exact_points.size());
    return a->exact_points.size() / 15;
}
// per node: 1 where a cavity walk stopped at the 100-step cap
void orc_ashape3d_capped_f32(void* h, uint8_t* out) {
    auto* a = static_cast<AlphaShape3D<float>*>(h);
    std::memcpy(out, a->node_capped.data(), a->n);
}
// sizes: [n, n_elems, n_edges, exact_calls, over_cap, broken_cycles]
void orc_ashape3d_sizes(void* h, int is_f64, uint64_t* out) {
    if (is_f64) {
        auto* a = (AlphaShape3D<double>*)h;
        out[0] = a->n; out[1] = a->n_elems; out[2] = a->n_edges; out[3] = a->exact_calls;
        out[4] = a->over_cap; out[5] = a->broken_cycles.load(); out[6] = a->capped_walks.load();
    } else {
        auto* a = (AlphaShape3D<float>*)h;
        out[0] = a->n; out[1] = a->n_elems; out[2] = a->n_edges; out[3] = a->exact_calls;
        out[4] = a->over_cap; out[5] = a->broken_cycles.load(); out[6] = a->capped_walks.load();
    }
}
void orc_ashape3d_get_f32(void* h, uint32_t* elem, uint32_t* row, uint32_t* neig, uint8_t* eb, uint8_t* nb,
                          uint32_t* perm, float* xyz_m, uint32_t* nineig) {
    get3d((AlphaShape3D<float>*)h, elem, row, neig, eb, nb, perm, xyz_m, nineig);
}
void orc_ashape3d_get_f64(void* h, uint32_t* elem, uint32_t* row, uint32_t* neig, uint8_t* eb, uint8_t* nb,
                          uint32_t* perm, double* xyz_m, uint32_t* nineig) {
    get3d((AlphaShape3D<double>*)h, elem, row, neig, eb, nb, perm, xyz_m, nineig);
}
void orc_ashape3d_free(void* h, int is_f64) {
    if (is_f64)
        delete (AlphaShape3D<double>*)h;
    else
        delete (AlphaShape3D<float>*)h;
}

int orc_lbvh2d_build_f32(uint32_t n, const float* xyr, int bits, int leaf_mode, uint64_t* codes,
                         uint64_t* codes_sorted, uint32_t* perm, int32_t* sep, uint8_t* cl, uint32_t* lp,
                         uint32_t* ip, float* ibox, float* gbox) {
    return lbvh_build<float, 2>(n, xyr, bits, leaf_mode, codes, codes_sorted, perm, sep, cl, lp, ip, ibox, gbox);
}
int orc_lbvh2d_build_f64(uint32_t n, const double* xyr, int bits, int leaf_mode, uint64_t* codes,
                         uint64_t* codes_sorted, uint32_t* perm, int32_t* sep, uint8_t* cl, uint32_t* lp,
                         uint32_t* ip, double* ibox, double* gbox) {
    return lbvh_build<double, 2>(n, xyr, bits, leaf_mode, codes, codes_sorted, perm, sep, cl, lp, ip, ibox, gbox);
}
int orc_lbvh3d_build_f64(uint32_t n, const double* xyr, int bits, int leaf_mode, uint64_t* codes,
                         uint64_t* codes_sorted, uint32_t* perm, int32_t* sep, uint8_t* cl, uint32_t* lp,
                         uint32_t* ip, double* ibox, double* gbox) {
    return lbvh_build<double, 3>(n, xyr, bits, leaf_mode, codes, codes_sorted, perm, sep, cl, lp, ip, ibox, gbox);
}
uint64_t orc_lbvh2d_collisions_f32(uint32_t n, const float* xyr, int bits) {
    return lbvh_collisions<float, 2>(n, xyr, bits);
}
uint64_t orc_lbvh2d_collisions_f64(uint32_t n, const double* xyr, int bits) {
    return lbvh_collisions<double, 2>(n, xyr, bits);
}

uint64_t orc_interleave(const uint32_t* q, int dim, int bits) { return interleave(q, dim, bits); }

float orc_orient2d_f32(const float* a, const float* b, const float* c) { return orient2d(a, b, c); }
double orc_orient2d_f64(const double* a, const double* b, const double* c) { return orient2d(a, b, c); }
int orc_orient2d_exact_f32(const float* a, const float* b, const float* c) { return exact::orient2d_sign(a, b, c); }
int orc_orient2d_exact_f64(const double* a, const double* b, const double* c) {
    return exact::orient2d_sign(a, b, c);
}
float orc_incircle_filter_f32(const float* a, const float* b, const float* c, const float* d, float* err) {
    return incircle_filter(a, b, c, d, err);
}
double orc_incircle_filter_f64(const double* a, const double* b, const double* c, const double* d, double* err) {
    return incircle_filter(a, b, c, d, err);
}
int orc_incircle_exact_f32(const float* a, const float* b, const float* c, const float* d) {
    return exact::incircle_sign(a, b, c, d);
}
int orc_incircle_exact_f64(const double* a, const double* b, const double* c, const double* d) {
    return exact::incircle_sign(a, b, c, d);
}

// ---- alpha-shape 2D: run, query sizes, copy outputs, free ----
void* orc_ashape2d_run_f32(uint32_t n, const float* xyr, int bits) { return ashape2d_run<float>(n, xyr, bits); }
void* orc_ashape2d_run_f64(uint32_t n, const double* xyr, int bits) { return ashape2d_run<double>(n, xyr, bits); }
void* orc_ashape2d_run_contract_f32(uint32_t n, const float* xyr, int bits, int contract) {
    return ashape2d_run<float>(n, xyr, bits, contract);
}

// sizes: [n, n_elems, n_edges, exact_calls]
void orc_ashape2d_sizes(void* h, int is_f64, uint64_t* out) {
    if (is_f64) {
        auto* a = (AlphaShape2D<double>*)h;
        out[0] = a->n; out[1] = a->n_elems; out[2] = a->n_edges; out[3] = a->exact_calls;
    } else {
        auto* a = (AlphaShape2D<float>*)h;
        out[0] = a->n; out[1] = a->n_elems; out[2] = a->n_edges; out[3] = a->exact_calls;
    }
}

void orc_ashape2d_get_f32(void* h, uint32_t* elem, uint32_t* row, uint32_t* neig, uint8_t* eb, uint8_t* nb,
                          uint32_t* perm, float* xy_m, uint32_t* nineig) {
    get2d((AlphaShape2D<float>*)h, elem, row, neig, eb, nb, perm, xy_m, nineig);
}
void orc_ashape2d_get_f64(void* h, uint32_t* elem, uint32_t* row, uint32_t* neig, uint8_t* eb, uint8_t* nb,
                          uint32_t* perm, double* xy_m, uint32_t* nineig) {
    get2d((AlphaShape2D<double>*)h, elem, row, neig, eb, nb, perm, xy_m, nineig);
}
void orc_ashape2d_free(void* h, int is_f64) {
    if (is_f64)
        delete (AlphaShape2D<double>*)h;
    else
        delete (AlphaShape2D<float>*)h;
}

// ---- P1 assembly: returns nnz; row/col/val/b must hold n+1 / cap / cap / n ----
int64_t orc_assemble_p1(uint32_t n, const double* xy, uint32_t ne, const uint32_t* elem, const uint8_t* bnd,
                        const double* f, uint64_t cap, uint32_t* row, uint32_t* col, double* val, double* b) {
    CsrD A;
    std::vector<double> bb;
    assemble_p1(n, xy, ne, elem, bnd, f, A, bb);
    if (A.col.size() > cap) return -(int64_t)A.col.size();
    std::memcpy(row, A.row.data(), sizeof(uint32_t) * (n + 1));
    std::memcpy(col, A.col.data(), sizeof(uint32_t) * A.col.size());
    std::memcpy(val, A.val.data(), sizeof(double) * A.val.size());
    std::memcpy(b, bb.data(), sizeof(double) * n);
    return (int64_t)A.col.size();
}

// ---- FP64 Jacobi-PCG with relative stopping ----
uint32_t orc_jacobi_pcg(uint32_t n, const uint32_t* row, const uint32_t* col, const double* val, const double* b,
                        double* x, double rtol, uint32_t maxit, double* true_rel) {
    CsrD A;
    A.n = n;
    A.row.assign(row, row + n + 1);
    A.col.assign(col, col + row[n]);
    A.val.assign(val, val + row[n]);
    return jacobi_pcg(A, b, x, rtol, maxit, true_rel);
}

// ---- timed CPU step (baseline): LBVH + alpha-shape, P1 assembly, and a
// bounded number of Jacobi-PCG iterations.  times_ms: [mesh, assembly,
// cg_total]; returns the CG iterations run.
uint32_t orc_step_f64(uint32_t n, const double* xyr, const double* f, uint32_t cg_iters, double* times_ms,
                      uint64_t* sizes) {
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    AlphaShape2D<double> a;
    a.morton_bits = 32;
    a.init(pack<double, 2>(n, xyr));
    a.compute();
    auto t1 = clk::now();
    std::vector<double> xy(2ull * n), fm(n);
    for (uint32_t i = 0; i < n; ++i) {
        xy[2 * i] = a.coords_m[i].c[0];
        xy[2 * i + 1] = a.coords_m[i].c[1];
        fm[i] = f[a.lbvh.perm[i]];
    }
    CsrD A;
    std::vector<double> bb;
    assemble_p1(n, xy.data(), a.n_elems, a.elem.data(), a.node_is_bnd.data(), fm.data(), A, bb);
    auto t2 = clk::now();
    std::vector<double> x(n, 0.0);
    double rel = 0;
    const uint32_t it = jacobi_pcg(A, bb.data(), x.data(), 1e-30, cg_iters, &rel);
    auto t3 = clk::now();
    times_ms[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    times_ms[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
    times_ms[2] = std::chrono::duration<double, std::milli>(t3 - t2).count();
    sizes[0] = a.n_elems;
    sizes[1] = A.col.size();
    return it;
}

// ---- timed CPU step solved to the tolerance (baseline), then an untimed
// reference solve: LBVH + alpha-shape, P1 assembly and FP64 Jacobi-PCG to
// a true relative residual of rtol from x = 0 (timed: times_ms [mesh,
// assembly, cg]); then the same PCG continued from that solution to
// ref_rtol (untimed), written to x_ref (Morton order).  info: [cg
// iterations, true rel after the timed solve, reference iterations, true
// rel of the reference]; sizes: [triangles, nnz].
void orc_step_solve_f64(uint32_t n, const double* xyr, const double* f, double rtol, double ref_rtol, uint32_t maxit,
                        double* times_ms, uint64_t* sizes, double* x_ref, double* info) {
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    AlphaShape2D<double> a;
    a.morton_bits = 32;
    a.init(pack<double, 2>(n, xyr));
    a.compute();
    auto t1 = clk::now();
    std::vector<double> xy(2ull * n), fm(n);
    for (uint32_t i = 0; i < n; ++i) {
        xy[2 * i] = a.coords_m[i].c[0];
        xy[2 * i + 1] = a.coords_m[i].c[1];
        fm[i] = f[a.lbvh.perm[i]];
    }
    CsrD A;
    std::vector<double> bb;
    assemble_p1(n, xy.data(), a.n_elems, a.elem.data(), a.node_is_bnd.data(), fm.data(), A, bb);
    auto t2 = clk::now();
    for (uint32_t i = 0; i < n; ++i) x_ref[i] = 0.0;
    double rel = 0;
    const uint32_t it = jacobi_pcg(A, bb.data(), x_ref, rtol, maxit, &rel);
    auto t3 = clk::now();
    times_ms[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    times_ms[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
    times_ms[2] = std::chrono::duration<double, std::milli>(t3 - t2).count();
    sizes[0] = a.n_elems;
    sizes[1] = A.col.size();
    double rel2 = 0;
    const uint32_t it2 = jacobi_pcg(A, bb.data(), x_ref, ref_rtol, maxit, &rel2);
    info[0] = it;
    info[1] = rel;
    info[2] = it2;
    info[3] = rel2;
}


// The oracle's squared norm and 2D cross product of the alpha filter
// (fmac::sumsq / fmac::diff in their plain mode), for comparison with the
// reference's Vec2f::sqnorm / cross (vec.hpp:216-238).
float orc_sumsq2_f32(float x, float y) {
    const float v[2] = {x, y};
    return orc::fmac::sumsq(0, false, v, 2);
}
float orc_cross2_f32(float ux, float uy, float vx, float vy) { return orc::fmac::diff(0, false, ux, vy, uy, vx); }

}  // extern "C"
