// Karras LBVH on the GPU (include/geometry/lbvh.hpp:61-362).
//
// Build sequence, all on the context stream with no host synchronisation:
//   1. box_partials   : per-block min/max of the sphere boxes   (lbvh.hpp:141-175)
//   2. morton_hist    : every block folds the partials into the global box,
//                       quantises its centres, interleaves the Morton code
//                       (lbvh.hpp:178-213) and accumulates the 8 digit
//                       histograms of the radix sort in LDS
//   3. radix passes   : onesweep, stable                      (lbvh.hpp:216-235)
//   4. gather         : objects into Morton order             (lbvh.hpp:237-242)
//   5. karras         : split search, child/leaf flags, parents (lbvh.hpp:246-308)
//   6. refit          : leaf-driven climb with an atomic arrival counter;
//                       the second arrival combines the sibling (lbvh.hpp:311-351)
//
// Layout in HBM: internal node i is one u32 word, sep | child_is_leaf << 30
// (sep < 2^30); internal boxes are BoxD<S,D> records so the two children of
// a node (indices sep and sep+1) are adjacent; parents carry the
// right-child bit in bit 31 as in the reference.
#pragma once
#include "geometry.hpp"
#include "primitives.hpp"

namespace stream_amd {

constexpr uint32_t kLeafCentre = 0;  // alpha-shape functor: box of the centre
constexpr uint32_t kLeafSphere = 1;  // ComputeBB_Functor: box of the sphere

template <class S, int D> struct LbvhGpu {
    static constexpr int kBoxBlocks = 512;

    uint32_t n = 0;
    int bits = 64 / D;
    uint32_t leaf_mode = kLeafCentre;

    DBuf<Sph4<S>> obj_m;             // Morton-ordered objects (+ extra slots)
    DBuf<BoxD<S, D>> gbox;           // [1] global box of the sphere boxes
    DBuf<BoxD<S, D>> partials;       // [kBoxBlocks]
    DBuf<uint64_t> k0, k1;
    DBuf<uint32_t> v0, v1;
    DBuf<uint32_t> sortws;
    const uint64_t* codes_sorted = nullptr;
    const uint32_t* perm = nullptr;  // map_sorted: Morton slot -> input index
    DBuf<uint32_t> node;             // [n-1]  sep | cl << 30
    DBuf<BoxD<S, D>> ibox;           // [n-1]
    DBuf<uint32_t> leaf_parent;      // [n]
    DBuf<uint32_t> internal_parent;  // [n-1]
    DBuf<uint32_t> touched;          // [n-1]

    // Build over n device objects.  extra_slots reserves room after the n
    // Morton-ordered objects (the alpha-shape's infinity nodes).
    void build(const Sph4<S>* d_obj, uint32_t n, int bits, uint32_t leaf_mode, uint32_t extra_slots,
               hipStream_t s);
    // Only steps 1-2: unsorted Morton codes into k0 (identity map in v0).
    void codes_only(const Sph4<S>* d_obj, uint32_t n, int bits, hipStream_t s);
    // Steps 1-4 (global box, codes, stable sort, gather into obj_m), shared
    // by the agglomerative build.
    void sort_only(const Sph4<S>* d_obj, uint32_t n, int bits, hipStream_t s);
};

}  // namespace stream_amd
