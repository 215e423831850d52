// Agglomerative LBVH (Apetrei 2014) device structure, shared by the TRBVH
// and the full 2D Delaunay mesher.  Replaces include/geometry/lbvh_agglo.hpp
// (LBVHa, :62-361).
//
// Leaves are nodes 0..n-1 (Morton order), internal nodes n..2n-2.
// children[2k], children[2k+1] are the left/right children of internal node
// k + n, range[2k..2k+1] its first/last leaf, split[k] the last leaf of its
// left child; parent_of[root] = 2n.
#pragma once
#include "lbvh.hpp"

namespace stream_amd {

template <class S, int D> struct Agglo {
    LbvhGpu<S, D> base;     // Morton codes, sort, Morton-ordered objects, global box
    DBuf<BoxD<S, D>> data;  // [2n-1] node boxes
    DBuf<uint64_t> delta;   // [n-1]
    DBuf<uint32_t> children, range, parent_of, split, touched, root;

    // Build over n >= 2 device objects.  sah / subtree (both or neither,
    // [2n-1]) receive the TRBVH costs and subtree object counts.
    void build(const Sph4<S>* d_obj, uint32_t n, int bits, uint32_t leaf_mode, hipStream_t s, float* sah = nullptr,
               uint32_t* subtree = nullptr);
};

}  // namespace stream_amd
