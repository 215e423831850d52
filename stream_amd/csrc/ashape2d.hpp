// 2D alpha-shape on the GPU (src/mesh/DirectAlphaShape2D.cpp:121-689).
//
// Pipeline (all on the context stream):
//   init    : LBVH over the spheres (centre boxes), then the three infinity
//             nodes appended after the Morton-ordered points (cpp:138-163).
//   compute : fill     - BVH traversal at radius 2*delta*alpha_i, candidate
//                        ids written to a fixed-capacity slab (cpp:310-357)
//             delaunay - per-node Bowyer-Watson insertion over the
//                        candidates with filtered predicates and the alpha
//                        filter (cpp:373-600).  A node whose filter cannot
//                        certify a sign is queued, not resolved in place.
//             exact    - the queued nodes redone with the exact stage and
//                        the simplified simulation of simplicity.
//             scans    - element and edge offsets (cpp:604-636)
//             compress - slab -> CSR edges + globally indexed triangles
//                        (cpp:642-689)
//
// Slab layout (replaces the reference's per-32-row hybrid ELL/CSR, whose
// block sizes come from a counting traversal and a host read-back): each
// node owns kSlabCap slots; slot j of node t lives at
//   ((t / 64) * kSlabCap + j) * 64 + t % 64,
// so the 64 lanes of a wave touching slot j of their nodes hit 256
// contiguous bytes.  With 288 GB of HBM the fixed capacity (1 KiB of
// candidates per node) removes the counting pass and every host sync from
// the meshing path.  kSlabCap = 256 also bounds local ids to u8, as in the
// reference (LocalElem uses uint8_t, DirectAlphaShape2D.hpp:80).
#pragma once
#include "lbvh.hpp"

namespace stream_amd {

constexpr uint32_t kSlabCap = 256;
constexpr uint32_t kNInf2D = 3;
constexpr uint32_t kMaxFill2D = 250;  // cpp:351
constexpr uint32_t kExactWaves = 4096;  // wavefronts of the exact redo (one queued node each at a time)

__host__ __device__ inline size_t slab_index(uint32_t t, uint32_t j) {
    return ((size_t)(t >> 6) * kSlabCap + j) * 64 + (t & 63u);
}
inline size_t slab_words(uint32_t n) { return (size_t)((n + 63) / 64) * kSlabCap * 64; }

// Directed edges reserved per node for the compressed output (the planar
// bound is 6; Stream_tune("edge_capacity", k) lowers it, which makes the
// capacity checks observable on small inputs).
uint32_t& ashape2d_edge_capacity();

template <class S> struct Ash2D {
    uint32_t n = 0;
    int bits = 32;
    DBuf<Sph4<S>> coords;  // input order
    LbvhGpu<S, 2> lbvh;    // lbvh.obj_m: Morton order + 3 infinity nodes
    DBuf<uint32_t> cand;   // slab: candidate global ids
    DBuf<uint16_t> tri;    // slab: local triangles (a | b << 8)
    DBuf<uint8_t> active;  // slab: per-candidate kept-triangle count
    DBuf<uint32_t> nineig, nfneig, row, elemrow;
    DBuf<uint8_t> node_is_bnd, nelem, nelem_out;
    DBuf<uint32_t> neig, elemglob;
    DBuf<uint8_t> edge_is_bnd;
    DBuf<uint32_t> queue;     // [0] = count, then node ids
    DBuf<uint64_t> scan_ws;
    uint32_t cap_edges = 0, cap_elems = 0;
    bool inited = false, computed = false;

    void set_nodes_host(uint32_t n, const S* xy, const S* alpha);
    void set_nodes_device(uint32_t n, const Sph4<S>* d_spheres);
    void init(hipStream_t s);
    void compute(hipStream_t s);
    uint32_t n_elems();  // synchronises
    uint32_t n_edges();  // synchronises
    uint32_t n_exact_nodes();  // synchronises
};

}  // namespace stream_amd
