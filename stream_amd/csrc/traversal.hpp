// Fixed-capacity traversal containers for BVH walks in device code
// (include/geometry/traversal_stack.hpp:31-110): a DFS stack and a binary
// min-heap keyed by a cost, both in registers / private memory.
//
// The heap keeps the reference's sift-up / sift-down order (a parent swaps
// with a child only on a strict cost comparison, the left child wins ties
// between children), so items of equal cost come out in the same order.
// push() on a full container returns false instead of writing past the end.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace stream_amd {

template <class IdxT, class CostT, int N> struct TraversalMinHeap {
    static constexpr int kMaxSize = N;
    struct Pair {
        IdxT first;
        CostT second;
    };
    int len = 0;
    Pair item[N];

    __host__ __device__ Pair peek() const { return item[0]; }
    __host__ __device__ bool empty() const { return len == 0; }

    __host__ __device__ bool push(IdxT idx, CostT cost) {
        if (len >= N) return false;
        item[len++] = {idx, cost};
        int i = len - 1;
        int p = (i - 1) >> 1;
        while (i > 0 && item[p].second > item[i].second) {
            const Pair t = item[i];
            item[i] = item[p];
            item[p] = t;
            i = p;
            p = (i - 1) >> 1;
        }
        return true;
    }

    __host__ __device__ void pop() {
        if (len == 0) return;
        item[0] = item[--len];
        int i = 0;
        while (true) {
            const int l = 2 * i + 1, r = 2 * i + 2;
            int s = i;
            if (l < len && item[l].second < item[s].second) s = l;
            if (r < len && item[r].second < item[s].second) s = r;
            if (s == i) break;
            const Pair t = item[i];
            item[i] = item[s];
            item[s] = t;
            i = s;
        }
    }
};

template <class IdxT, int N> struct TraversalStack {
    static constexpr int kMaxSize = N;
    int len = 0;
    IdxT item[N];

    __host__ __device__ IdxT peek() const { return item[len - 1]; }
    __host__ __device__ bool empty() const { return len == 0; }
    __host__ __device__ bool push(IdxT idx) {
        if (len >= N) return false;
        item[len++] = idx;
        return true;
    }
    __host__ __device__ void pop() {
        if (len) --len;
    }
};

}  // namespace stream_amd
