// Host evaluation of the device exact stage (predicates.hpp): the same
// expansion arithmetic, compiled for the host, for verification against
// exact rational arithmetic without a GPU.
#include <vector>

#include "predicates.hpp"
#include "runtime.hpp"

namespace stream_amd {
namespace {
template <class S> int exact_sign(int kind, const S* p, bool* over) {
    std::vector<double> ws(4096);
    const int cap = (int)ws.size();
    switch (kind) {
        case 0: return orient2d_exact(p, p + 2, p + 4, ws.data());
        case 1: return incircle_exact(p, p + 2, p + 4, p + 6, ws.data());
        case 2: return orient3d_exact(p, p + 3, p + 6, p + 9, ws.data(), cap, over);
        case 3: return insphere_exact(p, p + 3, p + 6, p + 9, p + 12, ws.data(), cap, over);
        default: throw std::invalid_argument("kind must be 0..3");
    }
}
}  // namespace
}  // namespace stream_amd

extern "C" {

// Exact sign of a predicate by the device exact stage run on the host:
// kind 0 orient2d (3 points x 2), 1 incircle (4 x 2), 2 orient3d (4 x 3),
// 3 insphere (5 x 3); pts float (fp64 = 0) or double.  The signs are those
// of the determinants as the device exact stage expands them (orient2d /
// orient3d / incircle / insphere of include/geometry/predicates.hpp).
int32_t Stream_exact_sign_host(int32_t kind, int32_t fp64, const void* pts, int32_t* sign) {
    return stream_amd::guarded([&] {
        if (!pts || !sign) throw std::invalid_argument("Stream_exact_sign_host: null argument");
        bool over = false;
        *sign = fp64 ? stream_amd::exact_sign<double>(kind, (const double*)pts, &over)
                     : stream_amd::exact_sign<float>(kind, (const float*)pts, &over);
        if (over) throw std::runtime_error("exact stage workspace exhausted");
    });
}

}  // extern "C"
