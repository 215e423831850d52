// 3D alpha-shape on the GPU: LBVH candidate search, local Bowyer-Watson
// tetrahedralisation of each node's star with filtered + exact insphere /
// orient3d predicates, the alpha filter, and compression to CSR.
//
// Replaces src/mesh/DirectAlphaShape3D.cpp:121-743 (AlphaShape3D::init,
// compute, compress).  The per-node work follows the reference statement by
// statement (candidate DFS order, insertion order, swap-with-last removal,
// the cavity-boundary cycle walk of Ray et al. with its 100-step cap, the
// re-triangulation order, the circumradius filter and the lowest-id owner),
// so the outputs are the same arrays as the CPU restatement in
// oracle/oracle3d.hpp.
//
// Layout: fixed-capacity slabs interleaved by wavefront, slot j of node t at
// ((t/64)*CAP + j)*64 + t%64: candidates (u32, CAP 124 = the reference's 128
// minus the four infinity nodes), local tetrahedra (u32 packing three u8
// local ids, CAP 256 >= 2*124 + 4 for a star of 124 link vertices) and
// per-candidate counts (u8).  The cavity cycle lives in LDS, 128 bytes per
// lane, as the reference's.
#pragma once
#include "lbvh.hpp"

namespace stream_amd {

constexpr uint32_t kNInf3D = 4;
constexpr uint32_t kCand3D = 124;    // real candidates per node
constexpr uint32_t kTet3D = 256;     // local tetrahedra per node
constexpr uint32_t kCycle3D = 128;   // local ids (4 infinity + 124)
constexpr uint32_t kExactWaves3D = 4096;  // wavefronts of the exact redo (one queued node each at a time)
constexpr uint32_t kExactWs3D = 512;  // doubles per lane (expansion components)

__host__ __device__ inline size_t slab3(uint32_t t, uint32_t j, uint32_t cap) {
    return ((size_t)(t >> 6) * cap + j) * 64 + (t & 63u);
}
inline size_t slab3_words(uint32_t n, uint32_t cap) { return (size_t)((n + 63) / 64) * cap * 64; }

template <class S> struct Ash3D {
    uint32_t n = 0;
    int bits = 21;
    DBuf<Sph4<S>> coords;  // input order
    LbvhGpu<S, 3> lbvh;    // lbvh.obj_m: Morton order + 4 infinity nodes
    DBuf<uint32_t> cand;   // slab (kCand3D)
    DBuf<uint32_t> tet;    // slab (kTet3D): a | b << 8 | c << 16
    DBuf<uint8_t> active;  // slab (kCand3D)
    DBuf<uint32_t> nineig, nfneig, row, elemrow;
    DBuf<uint8_t> node_is_bnd, nelem, nelem_out;
    DBuf<uint32_t> neig, elemglob;
    DBuf<uint8_t> edge_is_bnd;
    DBuf<uint32_t> queue;   // [0] = count, [1] = candidate overflows, then node ids
    DBuf<uint64_t> scan_ws;
    uint32_t n_elems_ = 0, n_edges_ = 0;
    bool inited = false, computed = false;

    void set_nodes_host(uint32_t n, const S* xyz, const S* alpha);
    void set_nodes_device(uint32_t n, const Sph4<S>* d_spheres);
    void init(hipStream_t s);
    void compute(hipStream_t s);  // compute + compress; synchronises once (output sizes)
    uint32_t n_elems() const { return computed ? n_elems_ : 0u; }
    uint32_t n_edges() const { return computed ? n_edges_ : 0u; }
    uint32_t n_exact_nodes();  // synchronises
    uint32_t n_over_cap();     // nodes with more than 124 candidates (clamped); synchronises
};

}  // namespace stream_amd
