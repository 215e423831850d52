// Thin C exports around the reference's self-contained geometry headers
// (include/geometry/predicates.hpp, morton.hpp, vec.hpp), compiled from
// where they lie under the reference tree by oracle/Makefile into
// oracle/_ref/libref_geom.so.  TEST INFRASTRUCTURE ONLY: it exists so the
// oracle's predicates and Morton encoders can be checked against the
// reference's own code (fp_tt = float, as the reference defines it).
#include <cstdint>

#include "defines.h"
#include "morton.hpp"
#include "predicates.hpp"
#include "vec.hpp"

extern "C" {

float ref_orient2d(const float* a, const float* b, const float* c) { return orient2d(a, b, c); }
float ref_orient2dexact(const float* a, const float* b, const float* c) { return orient2dexact(a, b, c); }
float ref_incircle(const float* a, const float* b, const float* c, const float* d, float* err) {
    return incircle(a, b, c, d, err);
}
float ref_incircleexact(const float* a, const float* b, const float* c, const float* d) {
    return incircleexact(a, b, c, d);
}
float ref_orient3d(const float* a, const float* b, const float* c, const float* d) { return orient3d(a, b, c, d); }
float ref_orient3dexact(const float* a, const float* b, const float* c, const float* d) {
    return orient3dexact(a, b, c, d);
}
float ref_insphere(const float* a, const float* b, const float* c, const float* d, const float* e, float* err) {
    return insphere(a, b, c, d, e, err);
}
float ref_insphereexact(const float* a, const float* b, const float* c, const float* d, const float* e) {
    return insphereexact(a, b, c, d, e);
}
uint64_t ref_encode2d(uint32_t x, uint32_t y) { return stream::geo::encode_magicbits2D(x, y); }
uint64_t ref_encode3d(uint32_t x, uint32_t y, uint32_t z) { return stream::geo::encode_magicbits3D(x, y, z); }
void ref_decode2d(uint64_t m, uint32_t* x, uint32_t* y) { stream::geo::decode_magicbits2D(m, *x, *y); }

// Vec2f sqnorm / cross as the reference evaluates them (vec.hpp:216-238).
float ref_vec2_sqnorm(float x, float y) {
    stream::geo::Vec<float, 2> v;
    v[0] = x;
    v[1] = y;
    return v.sqnorm();
}
float ref_vec2_cross(float ax, float ay, float bx, float by) {
    stream::geo::Vec<float, 2> a, b;
    a[0] = ax; a[1] = ay; b[0] = bx; b[1] = by;
    return a.cross(b);
}
}
