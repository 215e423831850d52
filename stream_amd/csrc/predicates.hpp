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
//     grow step, renormalised whenever it reaches kCompressAt components.
//     The expansion lives in a caller-provided per-lane slice
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

// Expansion workspace of the exact 2D stage in doubles (incircle: 48
// monomials of degree 4, at most 8 components each; renormalised at
// kCompressAt components, so far fewer are ever held).
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
// Expansion length at which the accumulator renormalises itself.  Growing
// one component at a time keeps the expansion nonoverlapping but not short:
// without renormalisation its length tracks the number of components added
// (hundreds for an insphere determinant), and every further add walks all of
// them.  Renormalised, it holds about as many components as its bit span
// needs (a handful for FP32 inputs).
constexpr int kCompressAt = 24;

struct ExactAcc {
    double* h;  // per-lane workspace
    int n;
    int cap = 1 << 30;  // workspace length (components)
    bool over = false;  // a component would not fit: the sign is unreliable

    // Shewchuk's renormalisation (COMPRESS, "Adaptive Precision Floating-
    // Point Arithmetic and Fast Robust Geometric Predicates", 1997), in
    // place: a downward pass merges each component into a running sum and
    // keeps the roundoff terms at the top of the array, an upward pass
    // merges them back from the top, keeping the nonzero roundoffs at the
    // bottom; the result is a nonoverlapping expansion of the same value,
    // increasing magnitude, with no two adjacent components overlapping.
    __host__ __device__ __forceinline__ void compress() {
        if (n < 2) return;
        int bottom = n - 1;
        double Q = h[bottom];
        for (int i = n - 2; i >= 0; --i) {
            const double e = h[i];
            const double x = Q + e;  // |Q| >= |e|: fast two-sum
            const double y = e - (x - Q);
            if (y != 0.0) {
                h[bottom--] = x;
                Q = y;
            } else {
                Q = x;
            }
        }
        int top = 0;
        for (int i = bottom + 1; i < n; ++i) {
            const double hv = h[i];
            const double x = hv + Q;  // |hv| >= |Q|: fast two-sum
            const double y = Q - (x - hv);
            if (y != 0.0) h[top++] = y;
            Q = x;
        }
        h[top] = Q;
        n = top + 1;
    }

    __host__ __device__ __forceinline__ void add(double b) {
        if (b == 0.0) return;
        if (n >= kCompressAt) compress();
        if (n + 1 > cap) {
            over = true;
            return;
        }
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
    // sign * f0 * f1 * f2 (exact, up to four components)
    __host__ __device__ __forceinline__ void add_prod3(double f0, double f1, double f2, double sign) {
        const double a = sign * f0;
        const double p0 = a * f1;
        const double p1 = fma(a, f1, -p0);
        const double q0 = p0 * f2, q1 = fma(p0, f2, -q0);
        const double q2 = p1 * f2, q3 = fma(p1, f2, -q2);
        add(q0);
        add(q1);
        add(q2);
        add(q3);
    }
    // sign * f0 * f1 * f2 * f3 * f4 (exact, up to sixteen components)
    __host__ __device__ __forceinline__ void add_prod5(double f0, double f1, double f2, double f3, double f4,
                                              double sign) {
        double p[16], q[16];
        const double a = sign * f0;
        p[0] = a * f1;
        p[1] = fma(a, f1, -p[0]);
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
        for (int k = 0; k < 8; ++k) {
            q[2 * k] = p[k] * f4;
            q[2 * k + 1] = fma(p[k], f4, -q[2 * k]);
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) add(q[k]);
    }
    // sign * f0 * f1 (exact, two components)
    __host__ __device__ __forceinline__ void add_prod2(double f0, double f1, double sign) {
        const double a = sign * f0;
        const double x = a * f1;
        const double y = fma(a, f1, -x);
        add(y);
        add(x);
    }
    // sign * f0 * f1 * f2 * f3 (exact, up to eight components)
    __host__ __device__ __forceinline__ void add_prod4(double f0, double f1, double f2, double f3, double sign) {
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
    __host__ __device__ __forceinline__ int sign() const {
        for (int i = n - 1; i >= 0; --i)
            if (h[i] != 0.0) return h[i] > 0.0 ? 1 : -1;
        return 0;
    }
};

// Exact sign of det | ax ay 1 ; bx by 1 ; cx cy 1 |.
template <class S>
__host__ __device__ __noinline__ int orient2d_exact(const S* a, const S* b, const S* c, double* ws, int cap = 1 << 30,
                                                    bool* over = nullptr) {
    ExactAcc acc{ws, 0, cap};
    const double ax = a[0], ay = a[1], bx = b[0], by = b[1], cx = c[0], cy = c[1];
    acc.add_prod2(ax, by, 1.0);
    acc.add_prod2(bx, cy, 1.0);
    acc.add_prod2(cx, ay, 1.0);
    acc.add_prod2(ax, cy, -1.0);
    acc.add_prod2(bx, ay, -1.0);
    acc.add_prod2(cx, by, -1.0);
    if (over) *over |= acc.over;
    return acc.sign();
}

// Exact sign of incircle(a,b,c,d) = det | px py px^2+py^2 1 |, expanded on
// the lift column: sum_k (-1)^k lift_k M_k, M_k the (x, y, 1) minor of the
// three other rows.
template <class S>
__host__ __device__ __noinline__ int incircle_exact(const S* a, const S* b, const S* c, const S* d, double* ws,
                                                    int cap = 1 << 30, bool* over = nullptr) {
    const double P[4][2] = {{(double)a[0], (double)a[1]},
                            {(double)b[0], (double)b[1]},
                            {(double)c[0], (double)c[1]},
                            {(double)d[0], (double)d[1]}};
    ExactAcc acc{ws, 0, cap};
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
    if (over) *over |= acc.over;
    return acc.sign();
}

// ---------------------------------------------------------------------------
// 3D predicates (include/geometry/predicates.hpp:585-1100).
// ---------------------------------------------------------------------------
template <class S> struct PredBounds3;
template <> struct PredBounds3<float> {
    static constexpr float o3dA = (float)((7.0 + 56.0 * 5.9604644775390625e-08) * 5.9604644775390625e-08);
    static constexpr float ispA = (float)((16.0 + 224.0 * 5.9604644775390625e-08) * 5.9604644775390625e-08);
};
template <> struct PredBounds3<double> {
    static constexpr double o3dA = (7.0 + 56.0 * 1.1102230246251565e-16) * 1.1102230246251565e-16;
    static constexpr double ispA = (16.0 + 224.0 * 1.1102230246251565e-16) * 1.1102230246251565e-16;
};

// orient3d filter (predicates.hpp:661-714): the determinant of the
// differences to d; *certain = false when its sign is not certified.
template <class S>
__device__ __forceinline__ S orient3d_filter(const S* pa, const S* pb, const S* pc, const S* pd, bool* certain) {
    const S adx = pa[0] - pd[0], bdx = pb[0] - pd[0], cdx = pc[0] - pd[0];
    const S ady = pa[1] - pd[1], bdy = pb[1] - pd[1], cdy = pc[1] - pd[1];
    const S adz = pa[2] - pd[2], bdz = pb[2] - pd[2], cdz = pc[2] - pd[2];
    const S bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const S cdxady = cdx * ady, adxcdy = adx * cdy;
    const S adxbdy = adx * bdy, bdxady = bdx * ady;
    const S det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    const S permanent = (fabs(bdxcdy) + fabs(cdxbdy)) * fabs(adz) + (fabs(cdxady) + fabs(adxcdy)) * fabs(bdz) +
                        (fabs(adxbdy) + fabs(bdxady)) * fabs(cdz);
    const S errbound = PredBounds3<S>::o3dA * permanent;
    *certain = (det > errbound) || (-det > errbound);
    return det;
}

// insphere filter (predicates.hpp:985-1100): det and its error bound.
template <class S>
__device__ __forceinline__ S insphere_filter(const S* pa, const S* pb, const S* pc, const S* pd, const S* pe, S* err) {
    const S aex = pa[0] - pe[0], bex = pb[0] - pe[0], cex = pc[0] - pe[0], dex = pd[0] - pe[0];
    const S aey = pa[1] - pe[1], bey = pb[1] - pe[1], cey = pc[1] - pe[1], dey = pd[1] - pe[1];
    const S aez = pa[2] - pe[2], bez = pb[2] - pe[2], cez = pc[2] - pe[2], dez = pd[2] - pe[2];
    const S aexbey = aex * bey, bexaey = bex * aey, ab = aexbey - bexaey;
    const S bexcey = bex * cey, cexbey = cex * bey, bc = bexcey - cexbey;
    const S cexdey = cex * dey, dexcey = dex * cey, cd = cexdey - dexcey;
    const S dexaey = dex * aey, aexdey = aex * dey, da = dexaey - aexdey;
    const S aexcey = aex * cey, cexaey = cex * aey, ac = aexcey - cexaey;
    const S bexdey = bex * dey, dexbey = dex * bey, bd = bexdey - dexbey;
    const S abc = fma(aez, bc, fma(-bez, ac, cez * ab));
    const S bcd = fma(bez, cd, fma(-cez, bd, dez * bc));
    const S cda = fma(cez, da, fma(dez, ac, aez * cd));
    const S dab = fma(dez, ab, fma(aez, bd, bez * da));
    const S alift = fma(aex, aex, fma(aey, aey, aez * aez));
    const S blift = fma(bex, bex, fma(bey, bey, bez * bez));
    const S clift = fma(cex, cex, fma(cey, cey, cez * cez));
    const S dlift = fma(dex, dex, fma(dey, dey, dez * dez));
    const S det = fma(dlift, abc, fma(-clift, dab, fma(blift, cda, -alift * bcd)));
    const S az = fabs(aez), bz = fabs(bez), cz = fabs(cez), dz = fabs(dez);
    const S p_aexbey = fabs(aexbey), p_bexaey = fabs(bexaey), p_bexcey = fabs(bexcey), p_cexbey = fabs(cexbey),
            p_cexdey = fabs(cexdey), p_dexcey = fabs(dexcey), p_dexaey = fabs(dexaey), p_aexdey = fabs(aexdey),
            p_aexcey = fabs(aexcey), p_cexaey = fabs(cexaey), p_bexdey = fabs(bexdey), p_dexbey = fabs(dexbey);
    const S permanent =
        fma(alift, fma(bz, (p_cexdey + p_dexcey), fma(cz, (p_dexbey + p_bexdey), dz * (p_bexcey + p_cexbey))),
            fma(blift, fma(cz, (p_dexaey + p_aexdey), fma(dz, (p_aexcey + p_cexaey), az * (p_cexdey + p_dexcey))),
                fma(clift,
                    fma(dz, (p_aexbey + p_bexaey), fma(az, (p_bexdey + p_dexbey), bz * (p_dexaey + p_aexdey))),
                    dlift * (fma(az, (p_bexcey + p_cexbey), fma(bz, (p_cexaey + p_aexcey), cz * (p_aexbey + p_bexaey)))))));
    *err = PredBounds3<S>::ispA * permanent;
    return det;
}

constexpr int kPerm3[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {0, 2, 1}, {1, 0, 2}, {2, 1, 0}};

// Exact sign of det | x y z 1 | over a, b, c, d (= sign of orient3d):
// Laplace on the ones column, 24 monomials of degree 3.  *over: workspace
// exhausted.
template <class S>
__host__ __device__ __noinline__ int orient3d_exact(const S* a, const S* b, const S* c, const S* d, double* ws, int cap,
                                           bool* over) {
    const double P[4][3] = {{(double)a[0], (double)a[1], (double)a[2]}, {(double)b[0], (double)b[1], (double)b[2]},
                            {(double)c[0], (double)c[1], (double)c[2]}, {(double)d[0], (double)d[1], (double)d[2]}};
    ExactAcc acc{ws, 0, cap};
    for (int k = 0; k < 4; ++k) {
        int r[3], m = 0;
        for (int q = 0; q < 4; ++q)
            if (q != k) r[m++] = q;
        const double cof = (k % 2 == 0) ? -1.0 : 1.0;
        for (int t = 0; t < 6; ++t)
            acc.add_prod3(P[r[kPerm3[t][0]]][0], P[r[kPerm3[t][1]]][1], P[r[kPerm3[t][2]]][2],
                          cof * (t < 3 ? 1.0 : -1.0));
    }
    *over |= acc.over;
    return acc.sign();
}

// Exact sign of det | x y z (x^2+y^2+z^2) 1 | over a..e (= sign of
// insphere): Laplace on the ones column, then on the lift column, the lift
// split into its three squares; 360 monomials of degree 5.
template <class S>
__host__ __device__ __noinline__ int insphere_exact(const S* a, const S* b, const S* c, const S* d, const S* e, double* ws,
                                           int cap, bool* over) {
    const double P[5][3] = {{(double)a[0], (double)a[1], (double)a[2]}, {(double)b[0], (double)b[1], (double)b[2]},
                            {(double)c[0], (double)c[1], (double)c[2]}, {(double)d[0], (double)d[1], (double)d[2]},
                            {(double)e[0], (double)e[1], (double)e[2]}};
    ExactAcc acc{ws, 0, cap};
    for (int k = 0; k < 5; ++k) {
        int R[4], m = 0;
        for (int q = 0; q < 5; ++q)
            if (q != k) R[m++] = q;
        const double ck = (k % 2 == 0) ? 1.0 : -1.0;
        for (int j = 0; j < 4; ++j) {
            int r[3], t3 = 0;
            for (int q = 0; q < 4; ++q)
                if (q != j) r[t3++] = R[q];
            const double cj = (j % 2 == 0) ? -1.0 : 1.0;
            for (int t = 0; t < 6; ++t) {
                const double st = t < 3 ? 1.0 : -1.0;
                for (int comp = 0; comp < 3; ++comp)
                    acc.add_prod5(P[R[j]][comp], P[R[j]][comp], P[r[kPerm3[t][0]]][0], P[r[kPerm3[t][1]]][1],
                                  P[r[kPerm3[t][2]]][2], ck * cj * st);
            }
        }
    }
    *over |= acc.over;
    return acc.sign();
}

}  // namespace stream_amd
