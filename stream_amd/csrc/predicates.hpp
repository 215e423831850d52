// Robust geometric predicates on the GPU (include/geometry/predicates.hpp).
//
// Two stages:
//   * the floating-point filters, evaluated in the input precision exactly
//     as the reference evaluates them (orient2d: predicates.hpp:385-418,
//     incircle with its fma form and error bound: predicates.hpp:544-583);
//     a filter either certifies the sign or reports "uncertain";
//   * an exact stage that returns the exact sign.  The determinant is
//     expanded into monomials, each monomial is formed exactly as a short
//     sum of doubles through error-free fma products, and all components are
//     accumulated into a nonoverlapping expansion with a zero-eliminating
//     grow step.  The expansion lives in a caller-provided per-lane slice
//     of global memory (kExactWsDoubles doubles), so the hot kernels keep a
//     small private footprint; only the deferred exact kernel passes one.
//
// Callers only use signs, so a certified filter value and the exact sign
// lead to identical decisions.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace stream_amd {

template <class S> struct PredEps;
template <> struct PredEps<float> {
    static constexpr double eps = 5.9604644775390625e-08;  // 2^-24
};
template <> struct PredEps<double> {
    static constexpr double eps = 1.1102230246251565404236316680908203125e-16;  // 2^-53
};
template <class S> struct PredBounds {
    static constexpr S ccwA = (S)((3.0 + 16.0 * PredEps<S>::eps) * PredEps<S>::eps);
    static constexpr S iccA = (S)((10.0 + 96.0 * PredEps<S>::eps) * PredEps<S>::eps);
};

// Doubles per lane needed by the exact 2D stage (incircle: 48 monomials of
// degree 4, at most 8 components each, plus the growing sum).
constexpr int kExactWsDoubles = 400;

// orient2d filter: returns det; *certain = false when the sign is not
// certified by ccwerrboundA.
template <class S>
__device__ __forceinline__ S orient2d_filter(const S* pa, const S* pb, const S* pc, bool* certain) {
    const S detleft = (pa[0] - pc[0]) * (pb[1] - pc[1]);
    const S detright = (pa[1] - pc[1]) * (pb[0] - pc[0]);
    const S det = detleft - detright;
    *certain = true;
    S detsum;
    if (detleft > (S)0) {
        if (detright <= (S)0) return det;
        detsum = detleft + detright;
    } else if (detleft < (S)0) {
        if (detright >= (S)0) return det;
        detsum = -detleft - detright;
    } else {
        return det;
    }
    const S errbound = PredBounds<S>::ccwA * detsum;
    if ((det >= errbound) || (-det >= errbound)) return det;
    *certain = false;
    return det;
}

// incircle filter: returns det and its error bound.
template <class S>
__device__ __forceinline__ S incircle_filter(const S* pa, const S* pb, const S* pc, const S* pd, S* err) {
    const S adx = pa[0] - pd[0], bdx = pb[0] - pd[0], cdx = pc[0] - pd[0];
    const S ady = pa[1] - pd[1], bdy = pb[1] - pd[1], cdy = pc[1] - pd[1];
    const S bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const S alift = fma(adx, adx, ady * ady);
    const S cdxady = cdx * ady, adxcdy = adx * cdy;
    const S blift = fma(bdx, bdx, bdy * bdy);
    const S adxbdy = adx * bdy, bdxady = bdx * ady;
    const S clift = fma(cdx, cdx, cdy * cdy);
    const S det = fma(alift, (bdxcdy - cdxbdy), fma(blift, (cdxady - adxcdy), clift * (adxbdy - bdxady)));
    const S perm = fma(alift, (fabs(bdxcdy) + fabs(cdxbdy)),
                       fma(blift, (fabs(cdxady) + fabs(adxcdy)), (fabs(adxbdy) + fabs(bdxady)) * clift));
    *err = PredBounds<S>::iccA * perm;
    return det;
}

// ---------------------------------------------------------------------------
// Exact stage.
// ---------------------------------------------------------------------------
struct ExactAcc {
    double* h;  // per-lane workspace
    int n;

    __device__ __forceinline__ void add(double b) {
        if (b == 0.0) return;
        double q = b;
        int k = 0;
        for (int i = 0; i < n; ++i) {
            const double hi = h[i];
            const double x = q + hi;
            const double bv = x - q;
            const double av = x - bv;
            const double y = (q - av) + (hi - bv);
            q = x;
            if (y != 0.0) h[k++] = y;
        }
        if (q != 0.0 || k == 0) h[k++] = q;
        n = k;
    }
    // sign * f0 * f1 (exact, two components)
    __device__ __forceinline__ void add_prod2(double f0, double f1, double sign) {
        const double a = sign * f0;
        const double x = a * f1;
        const double y = fma(a, f1, -x);
        add(y);
        add(x);
    }
    // sign * f0 * f1 * f2 * f3 (exact, up to eight components)
    __device__ __forceinline__ void add_prod4(double f0, double f1, double f2, double f3, double sign) {
        double p[8];
        const double a = sign * f0;
        p[0] = a * f1;
        p[1] = fma(a, f1, -p[0]);
        double q[8];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            q[2 * k] = p[k] * f2;
            q[2 * k + 1] = fma(p[k], f2, -q[2 * k]);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            p[2 * k] = q[k] * f3;
            p[2 * k + 1] = fma(q[k], f3, -p[2 * k]);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) add(p[k]);
    }
    __device__ __forceinline__ int sign() const {
        for (int i = n - 1; i >= 0; --i)
            if (h[i] != 0.0) return h[i] > 0.0 ? 1 : -1;
        return 0;
    }
};

// Exact sign of det | ax ay 1 ; bx by 1 ; cx cy 1 |.
template <class S> __device__ __noinline__ int orient2d_exact(const S* a, const S* b, const S* c, double* ws) {
    ExactAcc acc{ws, 0};
    const double ax = a[0], ay = a[1], bx = b[0], by = b[1], cx = c[0], cy = c[1];
    acc.add_prod2(ax, by, 1.0);
    acc.add_prod2(bx, cy, 1.0);
    acc.add_prod2(cx, ay, 1.0);
    acc.add_prod2(ax, cy, -1.0);
    acc.add_prod2(bx, ay, -1.0);
    acc.add_prod2(cx, by, -1.0);
    return acc.sign();
}

// Exact sign of incircle(a,b,c,d) = det | px py px^2+py^2 1 |, expanded on
// the lift column: sum_k (-1)^k lift_k M_k, M_k the (x, y, 1) minor of the
// three other rows.
template <class S>
__device__ __noinline__ int incircle_exact(const S* a, const S* b, const S* c, const S* d, double* ws) {
    const double P[4][2] = {{(double)a[0], (double)a[1]},
                            {(double)b[0], (double)b[1]},
                            {(double)c[0], (double)c[1]},
                            {(double)d[0], (double)d[1]}};
    ExactAcc acc{ws, 0};
    for (int k = 0; k < 4; ++k) {
        int r0, r1, r2;
        if (k == 0) { r0 = 1; r1 = 2; r2 = 3; }
        else if (k == 1) { r0 = 0; r1 = 2; r2 = 3; }
        else if (k == 2) { r0 = 0; r1 = 1; r2 = 3; }
        else { r0 = 0; r1 = 1; r2 = 2; }
        const double cof = (k % 2 == 0) ? 1.0 : -1.0;
        const double m[6][3] = {{P[r0][0], P[r1][1], 1.0},  {P[r1][0], P[r2][1], 1.0},
                                {P[r2][0], P[r0][1], 1.0},  {P[r0][0], P[r2][1], -1.0},
                                {P[r1][0], P[r0][1], -1.0}, {P[r2][0], P[r1][1], -1.0}};
        for (int t = 0; t < 6; ++t)
            for (int comp = 0; comp < 2; ++comp)
                acc.add_prod4(P[k][comp], P[k][comp], m[t][0], m[t][1], cof * m[t][2]);
    }
    return acc.sign();
}

}  // namespace stream_amd
