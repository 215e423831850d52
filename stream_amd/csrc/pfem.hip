// P1 pressure-Laplacian assembly on the alpha-shape, the mesh-and-solve
// step, and the reference numerics C-ABI (d_CSR / h_CSR / LinSys /
// PrecJacobi / SolverCG).
//
// P1 assembly has no counterpart in the reference (its README lists an
// fem module that does not exist).  Row i of the matrix is assembled by
// the thread of node i from its own kept local triangles, with no atomics:
//   K_ij = sum over the (at most two) triangles sharing edge ij of
//          -cot(theta_k) / 2, the cotangent evaluated from the triangle's
//          vertices in ascending id order so that rows i and j compute the
//          same bits (the matrix is bitwise symmetric);
//   K_ii = -(sum of the row's off-diagonals in ascending column order);
//   Dirichlet p = 0 on free-surface nodes: identity rows, zeroed columns;
//   b_i  = f_i * sum_T |T| / 3 (lumped load) on free rows.
// Columns are sorted; the row of node i starts at edge_row[i] + i.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ashape2d.hpp"
#include "mpnn.hpp"
#include "numerics.hpp"
#include "precond.hpp"

namespace stream_amd {

namespace {

__device__ __forceinline__ double cot_at(const double* pk, const double* pi, const double* pj) {
    const double ex = pi[0] - pk[0], ey = pi[1] - pk[1];
    const double fx = pj[0] - pk[0], fy = pj[1] - pk[1];
    const double dot = ex * fx + ey * fy;
    const double crs = fabs(ex * fy - ey * fx);
    return dot / crs;
}

struct P1Args {
    uint32_t n;
    const Sph4<double>* P;
    const uint32_t* cand;
    const uint16_t* tri;
    const uint8_t* nelem;
    const uint8_t* bnd;
    const uint32_t* erow;
    const uint32_t* eneig;
    const double* f;  // Morton-ordered source
    uint32_t* arow;
    uint32_t* acol;
    double* aval;
    double* idiag;
    float* idiag32;
    double* b;
    uint32_t cap;        // capacity of acol / aval
    uint32_t cap_edges;  // capacity of eneig (the mesh's edge list)
    uint32_t* flags;
};

// Rows of up to kP1Regs entries (the diagonal included) are built in
// registers: the columns kept sorted by a static-index insertion, each
// triangle's two contributions added by a predicated pass over the slots
// (no dynamic indexing, so nothing spills), and the row stored once.
// Longer rows take p1_row_global, which builds them in place in the
// output arrays.  Both give the same operations in the same order.
constexpr int kP1Regs = 16;

// Triangle t of node i: its two other vertices, the cotangent weights of
// the edges (i, ga) and (i, gb), and its area.
struct P1TriW {
    uint32_t ga, gb;
    double ka, kb, area;
};
__device__ __forceinline__ P1TriW p1_triangle(const P1Args& a, uint32_t i, uint32_t t) {
    const uint32_t n = a.n;
    auto gneig = [&](uint32_t j) -> uint32_t { return j < kNInf2D ? n + j : a.cand[slab_index(i, j - kNInf2D)]; };
    const uint16_t e = a.tri[slab_index(i, t)];
    P1TriW w;
    w.ga = gneig(e & 255u);
    w.gb = gneig(e >> 8);
    uint32_t v[3] = {i, w.ga, w.gb};
    if (v[1] < v[0]) { const uint32_t x = v[0]; v[0] = v[1]; v[1] = x; }
    if (v[2] < v[1]) { const uint32_t x = v[1]; v[1] = v[2]; v[2] = x; }
    if (v[1] < v[0]) { const uint32_t x = v[0]; v[0] = v[1]; v[1] = x; }
    double p[3][2];
    for (int k = 0; k < 3; ++k) {
        const Sph4<double> s = a.P[v[k]];
        p[k][0] = s.c[0];
        p[k][1] = s.c[1];
    }
    const double c0 = cot_at(p[0], p[1], p[2]);
    const double c1 = cot_at(p[1], p[0], p[2]);
    const double c2 = cot_at(p[2], p[0], p[1]);
    const double ux = p[1][0] - p[0][0], uy = p[1][1] - p[0][1];
    const double wx = p[2][0] - p[0][0], wy = p[2][1] - p[0][1];
    w.area = 0.5 * fabs(ux * wy - uy * wx);
    // k(edge) = -cot(opposite)/2: (v1,v2) <- c0, (v0,v2) <- c1, (v0,v1) <- c2;
    // the edge (i, ga) is opposite gb and (i, gb) opposite ga
    const double kk[3] = {-0.5 * c0, -0.5 * c1, -0.5 * c2};
    const int oa = (w.gb == v[0]) ? 0 : ((w.gb == v[1]) ? 1 : 2);
    const int ob = (w.ga == v[0]) ? 0 : ((w.ga == v[1]) ? 1 : 2);
    w.ka = kk[oa];
    w.kb = kk[ob];
    return w;
}

// Row i built in place in the output arrays (rows longer than kP1Regs).
__device__ void p1_row_global(const P1Args& a, uint32_t i, uint32_t e0, uint32_t nf, uint32_t base, double& mass) {
    uint32_t* col = a.acol + base;
    double* val = a.aval + base;
    col[0] = i;
    uint32_t cnt = 1;
    for (uint32_t k = 0; k < nf; ++k) {
        const uint32_t v = a.eneig[e0 + k];
        uint32_t p = cnt;
        while (p > 0 && col[p - 1] > v) {
            col[p] = col[p - 1];
            --p;
        }
        col[p] = v;
        ++cnt;
    }
    for (uint32_t k = 0; k < cnt; ++k) val[k] = 0.0;
    auto find = [&](uint32_t v) -> uint32_t {
        uint32_t lo = 0, hi = cnt;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (col[mid] <= v)
                lo = mid;
            else
                hi = mid;
        }
        return lo;
    };
    const uint32_t nt = a.nelem[i];
    for (uint32_t t = 0; t < nt; ++t) {
        const P1TriW w = p1_triangle(a, i, t);
        val[find(w.ga)] += w.ka;
        val[find(w.gb)] += w.kb;
        mass += w.area / 3.0;
    }
    const bool bi = a.bnd[i] != 0;
    double diag = 0.0;
    uint32_t dpos = 0;
    for (uint32_t k = 0; k < cnt; ++k) {
        if (col[k] == i) {
            dpos = k;
            continue;
        }
        diag -= val[k];
    }
    for (uint32_t k = 0; k < cnt; ++k)
        val[k] = (k == dpos) ? (bi ? 1.0 : diag) : ((bi || a.bnd[col[k]]) ? 0.0 : val[k]);
    const double dg = bi ? 1.0 : diag;
    a.idiag[i] = 1.0 / dg;
    a.idiag32[i] = (float)(1.0 / dg);
}

__global__ __launch_bounds__(256) void p1_assemble_kernel(P1Args a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t n = a.n;
    const uint32_t e0 = a.erow[i];
    const uint32_t nf = a.erow[i + 1] - e0;
    const uint32_t base = e0 + i;
    a.arow[i] = base;
    if (i == n - 1) a.arow[n] = a.erow[n] + n;
    if (e0 + nf > a.cap_edges || base + nf + 1 > a.cap) {
        // only reachable after an edge-capacity overflow of the mesh: the
        // row's edges (or its matrix entries) would lie past the end of
        // their arrays
        raise_flag(a.flags, kErrCapacity);
        return;
    }
    double mass = 0.0;
    const bool bi = a.bnd[i] != 0;
    if (nf + 1 > (uint32_t)kP1Regs) {
        p1_row_global(a, i, e0, nf, base, mass);
    } else {
        // sorted columns: insert each neighbour into the ascending list
        // (unused slots hold 0xffffffff, above every id)
        uint32_t c[kP1Regs];
        double v[kP1Regs];
#pragma unroll
        for (int k = 0; k < kP1Regs; ++k) {
            c[k] = 0xffffffffu;
            v[k] = 0.0;
        }
        c[0] = i;
        for (uint32_t q = 0; q < nf; ++q) {
            const uint32_t x = a.eneig[e0 + q];
#pragma unroll
            for (int k = kP1Regs - 1; k >= 1; --k) c[k] = (c[k] <= x) ? c[k] : (c[k - 1] <= x ? x : c[k - 1]);
            c[0] = (c[0] <= x) ? c[0] : x;
        }
        const uint32_t nt = a.nelem[i];
        for (uint32_t t = 0; t < nt; ++t) {
            const P1TriW w = p1_triangle(a, i, t);
#pragma unroll
            for (int k = 0; k < kP1Regs; ++k) {
                if (c[k] == w.ga) v[k] += w.ka;
                if (c[k] == w.gb) v[k] += w.kb;
            }
            mass += w.area / 3.0;
        }
        double diag = 0.0;
#pragma unroll
        for (int k = 0; k < kP1Regs; ++k)
            if (c[k] != i && c[k] != 0xffffffffu) diag -= v[k];
        const uint32_t cnt = nf + 1;
        uint32_t* col = a.acol + base;
        double* val = a.aval + base;
#pragma unroll
        for (int k = 0; k < kP1Regs; ++k) {
            if ((uint32_t)k < cnt) {
                const uint32_t ck = c[k];
                col[k] = ck;
                val[k] = (ck == i) ? (bi ? 1.0 : diag) : ((bi || a.bnd[ck]) ? 0.0 : v[k]);
            }
        }
        const double dg = bi ? 1.0 : diag;
        a.idiag[i] = 1.0 / dg;
        a.idiag32[i] = (float)(1.0 / dg);
    }
    a.b[i] = bi ? 0.0 : a.f[i] * mass;
}

template <class T>
__global__ __launch_bounds__(256) void permute_fwd_kernel(const T* __restrict__ in, const uint32_t* __restrict__ perm,
                                                          uint32_t n, T* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[perm[i]];
}
template <class T>
__global__ __launch_bounds__(256) void permute_bwd_kernel(const T* __restrict__ in, const uint32_t* __restrict__ perm,
                                                          uint32_t n, T* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[perm[i]] = in[i];
}

__global__ void pack_spheres_kernel(const double* __restrict__ xy, const double* __restrict__ alpha, uint32_t n,
                                    Sph4<double>* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Sph4<double> s;
    s.c[0] = xy[2ull * i];
    s.c[1] = xy[2ull * i + 1];
    s.c[2] = 0.0;
    s.r = alpha[i];
    out[i] = s;
}

// Fourier modes on the device: stream function psi = sum a sin(kx x + ky y
// + phi) (advection) and source f = sum (ax kx + ay ky) cos(kx x + ky y + phi).
struct Modes {
    uint32_t n = 0;
    double m[32][5];
};

// One forward-Euler step x += dt u(x), u = (dpsi/dy, -dpsi/dx).
__global__ __launch_bounds__(256) void advect_kernel(uint32_t n, Modes md, double dt, Sph4<double>* __restrict__ pts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Sph4<double> p = pts[i];
    double ux = 0.0, uy = 0.0;
    for (uint32_t k = 0; k < md.n; ++k) {
        const double kx = md.m[k][0], ky = md.m[k][1], ph = md.m[k][2], a = md.m[k][3];
        const double c = a * cos(kx * p.c[0] + ky * p.c[1] + ph);
        ux += ky * c;
        uy -= kx * c;
    }
    p.c[0] += dt * ux;
    p.c[1] += dt * uy;
    pts[i] = p;
}

__global__ __launch_bounds__(256) void source_kernel(uint32_t n, Modes md, const Sph4<double>* __restrict__ pts,
                                                     double* __restrict__ f) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Sph4<double> p = pts[i];
    double s = 0.0;
    for (uint32_t k = 0; k < md.n; ++k) {
        const double kx = md.m[k][0], ky = md.m[k][1], ph = md.m[k][2];
        s += (md.m[k][3] * kx + md.m[k][4] * ky) * cos(kx * p.c[0] + ky * p.c[1] + ph);
    }
    f[i] = s;
}

// ---------------------------------------------------------------------------
// Fractional-step (projection) velocity update on the P1 alpha-shape mesh,
// lumped mass.  With K_ij = sum_T |T| grad phi_i . grad phi_j (the assembled
// pressure matrix before its Dirichlet rows), m_i = sum_T |T| / 3,
// D_i(u) = sum_T (|T| / 3) div u_h|T and G_i(p) = sum_T (|T| / 3) grad p_h|T
// over the triangles T of node i:
//   predictor   u*_i = u_i + dt (g - nu (K u)_i / m_i),
//   pressure    K p = -(rho / dt) D(u*), p = 0 on free-surface nodes,
//   correction  u_i <- u*_i - (dt / rho) G_i(p) / m_i.
// A node without triangles (m_i = 0) keeps u*_i = u_i + dt g.  Each node
// sums its own triangles, the vertices of each in ascending id order, so
// the results do not depend on scheduling.
// ---------------------------------------------------------------------------
struct FlowArgs {
    uint32_t n;
    const Sph4<double>* P;  // Morton-ordered records
    const uint32_t* cand;
    const uint16_t* tri;
    const uint8_t* nelem;
    const uint8_t* bnd;
    const double2* u;   // velocity (Morton)
    double2* us;        // u*
    double* mass;
    const double* p;    // pressure (Morton)
    double2* u_new;
    double* b;          // pressure right-hand side
    double* div;        // D(u) of the field `u` (flow_div_kernel with div != null)
    double dt, rho, nu, gx, gy;
};

// Vertices (ascending ids), area and basis gradients of triangle (i, ga, gb).
struct P1Tri {
    uint32_t v[3];
    double area;
    double gx[3], gy[3];
};
__device__ __forceinline__ P1Tri p1_tri(const Sph4<double>* P, uint32_t i, uint32_t ga, uint32_t gb) {
    P1Tri t;
    t.v[0] = i;
    t.v[1] = ga;
    t.v[2] = gb;
    if (t.v[1] < t.v[0]) { const uint32_t x = t.v[0]; t.v[0] = t.v[1]; t.v[1] = x; }
    if (t.v[2] < t.v[1]) { const uint32_t x = t.v[1]; t.v[1] = t.v[2]; t.v[2] = x; }
    if (t.v[1] < t.v[0]) { const uint32_t x = t.v[0]; t.v[0] = t.v[1]; t.v[1] = x; }
    double x[3], y[3];
    for (int k = 0; k < 3; ++k) {
        const Sph4<double> s = P[t.v[k]];
        x[k] = s.c[0];
        y[k] = s.c[1];
    }
    const double a2 = (x[1] - x[0]) * (y[2] - y[0]) - (y[1] - y[0]) * (x[2] - x[0]);  // signed 2|T|
    t.area = 0.5 * fabs(a2);
    const double inv = 1.0 / a2;
    t.gx[0] = (y[1] - y[2]) * inv;
    t.gy[0] = (x[2] - x[1]) * inv;
    t.gx[1] = (y[2] - y[0]) * inv;
    t.gy[1] = (x[0] - x[2]) * inv;
    t.gx[2] = (y[0] - y[1]) * inv;
    t.gy[2] = (x[1] - x[0]) * inv;
    return t;
}

template <class F> __device__ __forceinline__ void for_node_triangles(const FlowArgs& a, uint32_t i, F&& f) {
    auto gneig = [&](uint32_t j) -> uint32_t { return j < kNInf2D ? a.n + j : a.cand[slab_index(i, j - kNInf2D)]; };
    const uint32_t nt = a.nelem[i];
    for (uint32_t t = 0; t < nt; ++t) {
        const uint16_t e = a.tri[slab_index(i, t)];
        f(p1_tri(a.P, i, gneig(e & 255u), gneig(e >> 8)));
    }
}

__device__ __forceinline__ int local_of(const P1Tri& t, uint32_t i) { return t.v[0] == i ? 0 : (t.v[1] == i ? 1 : 2); }

__global__ __launch_bounds__(256) void flow_predict_kernel(FlowArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double2 ui = a.u[i];
    double m = 0.0, kx = 0.0, ky = 0.0;
    for_node_triangles(a, i, [&](const P1Tri& t) {
        const int li = local_of(t, i);
        m += t.area / 3.0;
        for (int k = 0; k < 3; ++k) {
            if (k == li) continue;
            const double kij = t.area * (t.gx[li] * t.gx[k] + t.gy[li] * t.gy[k]);
            const double2 uk = a.u[t.v[k]];
            kx += kij * (uk.x - ui.x);  // rows of K sum to zero: (K u)_i = sum_j K_ij (u_j - u_i)
            ky += kij * (uk.y - ui.y);
        }
    });
    double2 us;
    us.x = ui.x + a.dt * a.gx;
    us.y = ui.y + a.dt * a.gy;
    if (m > 0.0) {
        us.x -= a.dt * a.nu * kx / m;
        us.y -= a.dt * a.nu * ky / m;
    }
    a.us[i] = us;
    a.mass[i] = m;
}

// b_i = -(rho / dt) D_i(u*) on free nodes, 0 on free-surface nodes; with
// a.div set, D_i(a.u) instead (diagnostics).
__global__ __launch_bounds__(256) void flow_div_kernel(FlowArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double2* u = a.div ? a.u : a.us;
    double d = 0.0;
    for_node_triangles(a, i, [&](const P1Tri& t) {
        double dv = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double2 uk = u[t.v[k]];
            dv += uk.x * t.gx[k] + uk.y * t.gy[k];
        }
        d += t.area / 3.0 * dv;
    });
    if (a.div)
        a.div[i] = d;
    else
        a.b[i] = a.bnd[i] ? 0.0 : -(a.rho / a.dt) * d;
}

__global__ __launch_bounds__(256) void flow_correct_kernel(FlowArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double gx = 0.0, gy = 0.0;
    for_node_triangles(a, i, [&](const P1Tri& t) {
        double px = 0.0, py = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double pk = a.p[t.v[k]];
            px += pk * t.gx[k];
            py += pk * t.gy[k];
        }
        gx += t.area / 3.0 * px;
        gy += t.area / 3.0 * py;
    });
    double2 u = a.us[i];
    const double m = a.mass[i];
    if (m > 0.0) {
        u.x -= (a.dt / a.rho) * gx / m;
        u.y -= (a.dt / a.rho) * gy / m;
    }
    a.u_new[i] = u;
}

// Lagrangian move of the particles: x += dt u (input order).
__global__ __launch_bounds__(256) void move_kernel(uint32_t n, double dt, const double2* __restrict__ u,
                                                   Sph4<double>* __restrict__ pts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Sph4<double> p = pts[i];
    const double2 v = u[i];
    p.c[0] += dt * v.x;
    p.c[1] += dt * v.y;
    pts[i] = p;
}

}  // namespace

// ---------------------------------------------------------------------------
// The PFEM mesh-and-solve step on device-resident points.
// ---------------------------------------------------------------------------
struct Step2D {
    Ash2D<double> mesh;
    DBuf<Sph4<double>> pts;  // input order
    DBuf<double> f_in, f_m, b, x, p_out;
    DevCsr<double> A;
    Sell<double> S64;  // FP64 iteration matrix (standard form)
    Sell<float> S32;   // FP32 iteration matrix (row-sum form)
    LowerFactor<float> L32;  // M^-1 = L L^T (FSAI or learned)
    DBuf<float> idiag32;
    DBuf<double> idiag;
    CgWork work;
    SolveStats last;
    double t_mesh = 0, t_asm = 0, t_prep = 0, t_solve = 0;
    bool solved = false;  // x holds the solution of the current system
    uint32_t n = 0;
    PrecLearned* learned = nullptr;  // network of solver 3 (not owned)
    bool learned_fallback = false;   // the last learned factor was rejected
    Modes source;                    // device source f(x) (n = 0: host-given f)
    // projection step: nodal velocity (input order), Morton copies, mass
    DBuf<double2> u_in, u_m, us_m, unew_m;
    DBuf<double> mass, div_m;
    bool has_velocity = false;
    double t_flow = 0;

    // Point the network's graph at this step's system (after assembly).
    void bind_graph(MpnnGraph& g, hipStream_t s, bool transpose = true) {
        g.n = n;
        g.nnz = A.nnz;
        g.row = A.row.p;
        g.col = A.col.p;
        g.val = A.val.p;
        g.xy = reinterpret_cast<const double*>(mesh.lbvh.obj_m.p);  // Sph4: c[0], c[1] first
        g.xy_stride = sizeof(Sph4<double>) / sizeof(double);
        g.dim = 2;
        g.bnd = mesh.node_is_bnd.p;
        mpnn_prepare(g, s, transpose);
    }

    void advect(const Modes& md, double dt, uint32_t steps) {
        hipStream_t s = ctx().stream;
        for (uint32_t k = 0; k < steps; ++k)
            hipLaunchKernelGGL(advect_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, md, dt, pts.p);
        if (source.n) update_source(s);
        SA_HIP(hipStreamSynchronize(s));
    }

    void update_source(hipStream_t s) {
        f_in.resize(n ? n : 1);
        hipLaunchKernelGGL(source_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, source, pts.p, f_in.p);
    }

    FlowArgs flow_args() {
        FlowArgs a;
        a.n = n;
        a.P = mesh.lbvh.obj_m.p;
        a.cand = mesh.cand.p;
        a.tri = mesh.tri.p;
        a.nelem = mesh.nelem.p;
        a.bnd = mesh.node_is_bnd.p;
        a.u = u_m.p;
        a.us = us_m.p;
        a.mass = mass.p;
        a.p = x.p;
        a.u_new = unew_m.p;
        a.b = b.p;
        a.div = nullptr;
        a.dt = a.rho = a.nu = a.gx = a.gy = 0.0;
        return a;
    }

    void set_velocity(const double* h_uv) {
        u_in.resize(n ? n : 1);
        h2d(u_in.p, reinterpret_cast<const double2*>(h_uv), n);
        has_velocity = true;
    }

    // One fractional step on the current points: mesh, assemble, predictor,
    // pressure right-hand side, solve (solver as in solve()), correction;
    // move = true then moves the points by dt u (the PFEM Lagrangian step).
    void projection_step(int bits, int solver, double dt, double rho, double nu, double gx, double gy, double rtol,
                         uint32_t maxit, bool move) {
        if (!has_velocity) throw std::invalid_argument("projection step needs a velocity (PfemStep2D_set_velocity)");
        if (!(dt > 0.0) || !(rho > 0.0) || !(nu >= 0.0)) throw std::invalid_argument("projection step: bad dt, rho or nu");
        mesh_and_assemble(bits, solver);
        hipStream_t s = ctx().stream;
        hipEvent_t e0, e1;
        SA_HIP(hipEventCreate(&e0));
        SA_HIP(hipEventCreate(&e1));
        SA_HIP(hipEventRecord(e0, s));
        u_m.resize(n);
        us_m.resize(n);
        unew_m.resize(n);
        mass.resize(n);
        permute_elements(u_in.p, u_m.p, mesh.lbvh.perm, n, sizeof(double2), true, s);
        FlowArgs a = flow_args();
        a.dt = dt;
        a.rho = rho;
        a.nu = nu;
        a.gx = gx;
        a.gy = gy;
        const dim3 grid(grid_for(n, 256)), blk(256);
        hipLaunchKernelGGL(flow_predict_kernel, grid, blk, 0, s, a);
        hipLaunchKernelGGL(flow_div_kernel, grid, blk, 0, s, a);
        SA_HIP(hipEventRecord(e1, s));
        SA_HIP(hipEventSynchronize(e1));
        float t0 = 0;
        SA_HIP(hipEventElapsedTime(&t0, e0, e1));
        solve(solver, rtol, maxit);
        SA_HIP(hipEventRecord(e0, s));
        hipLaunchKernelGGL(flow_correct_kernel, grid, blk, 0, s, a);
        permute_elements(unew_m.p, u_in.p, mesh.lbvh.perm, n, sizeof(double2), false, s);
        if (move) {
            hipLaunchKernelGGL(move_kernel, grid, blk, 0, s, n, dt, u_in.p, pts.p);
            if (source.n) update_source(s);
        }
        SA_HIP(hipEventRecord(e1, s));
        SA_HIP(hipEventSynchronize(e1));
        float t1 = 0;
        SA_HIP(hipEventElapsedTime(&t1, e0, e1));
        t_flow = t0 + t1;
        SA_HIP(hipEventDestroy(e0));
        SA_HIP(hipEventDestroy(e1));
        if (move) {
            A.n = 0;  // the system and solution belong to the points before the move
            A.nnz = 0;
            solved = false;
        }
    }

    // Diagnostics of the last projection step (Morton order): u*, the new
    // velocity, the lumped mass and D(u_new).
    void flow_fields(double* h_us, double* h_unew, double* h_mass, double* h_div) {
        if (!unew_m.p || unew_m.n != n) throw std::logic_error("no projection step for the current points");
        hipStream_t s = ctx().stream;
        if (h_div) {
            div_m.resize(n);
            FlowArgs a = flow_args();
            a.u = unew_m.p;
            a.div = div_m.p;
            hipLaunchKernelGGL(flow_div_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, a);
            d2h(h_div, div_m.p, n);
        }
        if (h_us) d2h(reinterpret_cast<double2*>(h_us), us_m.p, n);
        if (h_unew) d2h(reinterpret_cast<double2*>(h_unew), unew_m.p, n);
        if (h_mass) d2h(h_mass, mass.p, n);
    }

    void set_points(uint32_t n_, const double* h_xy, const double* h_alpha, const double* h_f) {
        if (n_ != n) has_velocity = false;
        n = n_;
        A.n = 0;  // the assembled system and solution belong to the previous points
        A.nnz = 0;
        solved = false;
        hipStream_t s = ctx().stream;
        DBuf<double> dxy, da;
        dxy.resize(2ull * n);
        da.resize(n);
        h2d(dxy.p, h_xy, 2ull * n);
        h2d(da.p, h_alpha, n);
        pts.resize(n);
        hipLaunchKernelGGL(pack_spheres_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, dxy.p, da.p, n, pts.p);
        f_in.resize(n);
        h2d(f_in.p, h_f, n);
        SA_HIP(hipStreamSynchronize(s));
    }

    // mesh (LBVH + alpha-shape), permute the source, assemble, and build the
    // solver's matrix copy and preconditioner (solver < 0: none)
    void mesh_and_assemble(int bits, int solver) {
        hipStream_t s = ctx().stream;
        hipEvent_t ev[4];
        for (auto& e : ev) SA_HIP(hipEventCreate(&e));
        SA_HIP(hipEventRecord(ev[0], s));
        mesh.bits = bits;
        mesh.set_nodes_device(n, pts.p);
        mesh.init(s);
        mesh.compute(s);
        SA_HIP(hipEventRecord(ev[1], s));
        assemble(s);
        SA_HIP(hipEventRecord(ev[2], s));
        if (solver >= 0) prepare(solver, s);
        SA_HIP(hipEventRecord(ev[3], s));
        SA_HIP(hipEventSynchronize(ev[3]));
        float a = 0, c = 0, d = 0;
        SA_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
        SA_HIP(hipEventElapsedTime(&c, ev[1], ev[2]));
        SA_HIP(hipEventElapsedTime(&d, ev[2], ev[3]));
        t_mesh = a;
        t_asm = c;
        t_prep = d;
        for (auto& e : ev) SA_HIP(hipEventDestroy(e));
    }

    void assemble(hipStream_t s) {
        solved = false;
        const uint32_t cap = mesh.cap_edges + n;  // nnz <= edges + n
        A.n = n;
        A.row.resize(n + 1);
        A.col.resize(cap);
        A.val.resize(cap);
        idiag.resize(n);
        idiag32.resize(n);
        b.resize(n);
        f_m.resize(n);
        x.resize(n);
        hipLaunchKernelGGL((permute_fwd_kernel<double>), dim3(grid_for(n, 256)), dim3(256), 0, s, f_in.p,
                           mesh.lbvh.perm, n, f_m.p);
        P1Args a;
        a.n = n;
        a.P = mesh.lbvh.obj_m.p;
        a.cand = mesh.cand.p;
        a.tri = mesh.tri.p;
        a.nelem = mesh.nelem.p;
        a.bnd = mesh.node_is_bnd.p;
        a.erow = mesh.row.p;
        a.eneig = mesh.neig.p;
        a.f = f_m.p;
        a.arow = A.row.p;
        a.acol = A.col.p;
        a.aval = A.val.p;
        a.idiag = idiag.p;
        a.idiag32 = idiag32.p;
        a.b = b.p;
        a.cap = cap;
        a.cap_edges = mesh.cap_edges;
        a.flags = ctx().d_flags;
        section_mark(s, kMarkAsm0);
        hipLaunchKernelGGL(p1_assemble_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, a);
        section_mark(s, kMarkAsm);
        A.nnz = read_scalar(A.row.p + n);
        // a capacity overflow of the mesh (edges or triangles) or of the
        // matrix leaves entries unwritten or past the end of A.col / A.val:
        // stop before anything reads them
        const uint32_t fl = read_scalar(ctx().d_flags + kFlagError);
        if (fl || A.nnz > cap) {
            fill<uint32_t>(ctx().d_flags, 1, 0u, s);
            A.nnz = 0;
            throw std::runtime_error(fl == kErrCapacity || fl == 0
                                         ? "mesh or matrix capacity overflow (edges > 6n + 64 or triangles > 2n + 64)"
                                         : "device error flag raised during meshing or assembly");
        }
    }

    // SELL copy for the chosen solver (FP64 standard or FP32 row-sum form)
    // and its factor.  The learned factor is rejected (Jacobi is used
    // instead) when any of its values is non-finite or its diagonal is not
    // positive.
    void prepare(int solver, hipStream_t s) {
        learned_fallback = false;
        if (solver == 0) {
            build_sell<double>(n, A.row.p, A.col.p, A.val.p, nullptr, kSellStandard, kSellOffDiag, S64, s);
        } else {
            build_sell<float>(n, A.row.p, A.col.p, A.val.p, nullptr, kSellRowSum, kSellOffDiag, S32, s);
            if (solver == 2) fsai_factor<float>(A, L32, s);
            if (solver == 3) {
                if (!learned) throw std::invalid_argument("solver 3 needs a network (PfemStep2D_set_learned)");
                factor_structure<float>(A, L32, s);
                bind_graph(learned->graph, s, false);  // inference: no transpose map
                learned->net.forward(learned->graph, s, false);
                learned->net.pack(learned->graph, L32, s);
                const uint32_t bad = read_scalar(ctx().d_flags + kFlagLearned);
                if (bad) {
                    fill<uint32_t>(ctx().d_flags + kFlagLearned, 1, 0u, s);
                    learned_fallback = true;
                }
            }
        }
    }

    // solver: 0 = FP64 Jacobi-PCG, 1 = mixed FP32/FP64 Jacobi-PCG,
    // 2 = mixed PCG with the FSAI L L^T factor, 3 = mixed PCG with the
    // learned L L^T factor (Jacobi when the factor was rejected)
    void solve(int solver, double rtol, uint32_t maxit) {
        if (solver == 3 && learned_fallback) solver = 1;
        hipStream_t s = ctx().stream;
        fill<double>(x.p, n, 0.0, s);
        CgOptions o;
        o.rtol = rtol;
        o.maxit = maxit;
        o.batch = tuning().batch;
        if (solver == 0) {
            cg_solve<double>(A, S64, false, kPrecJacobi, idiag.p, nullptr, b.p, x.p, o, work, last, s);
        } else if (solver == 1) {
            o.replace_factor = tuning().replace_factor;
            cg_solve<float>(A, S32, true, kPrecJacobi, idiag32.p, nullptr, b.p, x.p, o, work, last, s);
        } else if (solver == 2 || solver == 3) {
            o.replace_factor = tuning().replace_factor;
            cg_solve<float>(A, S32, true, kPrecLLt, nullptr, &L32, b.p, x.p, o, work, last, s);
        } else {
            throw std::invalid_argument("unknown solver kind");
        }
        t_solve = last.ms;
        solved = true;
    }

    // the system of the current points (mesh + assembly done)
    void require_system() const {
        if (!A.nnz || A.n != n) throw std::logic_error("no assembled system for the current points (run a step first)");
    }
    void require_solution() const {
        require_system();
        if (!solved) throw std::logic_error("no solution for the current system (run a step first)");
    }

    void solution_input_order(double* h_p) {
        require_solution();
        hipStream_t s = ctx().stream;
        p_out.resize(n);
        hipLaunchKernelGGL((permute_bwd_kernel<double>), dim3(grid_for(n, 256)), dim3(256), 0, s, x.p,
                           mesh.lbvh.perm, n, p_out.p);
        d2h(h_p, p_out.p, n);
    }
};

}  // namespace stream_amd

using namespace stream_amd;

// ===========================================================================
// Reference numerics C-ABI (csr.hpp:51-68, linear_system.hpp:48-54,
// preconditioner.hpp:39-40, solver.hpp:54-61) plus FP64 extensions.
// ===========================================================================
struct h_CSR {
    uint32_t n = 0;
    std::vector<uint32_t> row, col;
    std::vector<float> val32;
    std::vector<double> val64;
    bool f64 = false;
};
struct d_CSR {
    DevCsr<double> A;  // row/col always; val when f64
    DBuf<float> val32;
    bool f64 = false;
    Sell<float> S32;   // standard form, from the FP32 values
    Sell<double> S64;  // standard form, from the FP64 values
    bool s32_ok = false, s64_ok = false;
    const Sell<float>& sell32() {
        if (!s32_ok) {
            build_sell<float>(A.n, A.row.p, A.col.p, nullptr, val32.p, kSellStandard, kSellOffDiag, S32,
                              ctx().stream);
            s32_ok = true;
        }
        return S32;
    }
    const Sell<double>& sell64() {
        ensure_f64();
        if (!s64_ok) {
            build_sell<double>(A.n, A.row.p, A.col.p, A.val.p, nullptr, kSellStandard, kSellOffDiag, S64,
                               ctx().stream);
            s64_ok = true;
        }
        return S64;
    }
    void ensure_f64() {
        if (f64) return;
        // widen the FP32 values exactly
        std::vector<float> h(A.nnz);
        d2h(h.data(), val32.p, A.nnz);
        std::vector<double> d(h.begin(), h.end());
        A.val.resize(A.nnz ? A.nnz : 1);
        h2d(A.val.p, d.data(), A.nnz);
        f64 = true;
    }
};
struct LinSys {
    uint32_t n = 0;
    const d_CSR* A = nullptr;  // shares the CSR, as LinSys_set copies the DeviceCSR handle
    DBuf<double> b;
};
struct PrecJacobi_st {
    DBuf<double> idiag;
    DBuf<float> idiag32;
    uint32_t missing = 0;
};
struct SolverCG {
    CgWork work;
    SolveStats last;
    DBuf<double> x;
};

namespace {
std::vector<float> h_vec_f32(const h_CSR* h) {
    if (!h->f64) return h->val32;
    return std::vector<float>(h->val64.begin(), h->val64.end());
}

__global__ void jacobi_kernel(const uint32_t* row, const uint32_t* col, const float* v32, const double* v64,
                              uint32_t n, double* idiag, float* idiag32, uint32_t* missing) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    bool found = false;
    for (uint32_t k = row[i]; k < row[i + 1]; ++k)
        if (col[k] == i) {
            d = v64 ? v64[k] : (double)v32[k];
            found = true;
            break;
        }
    if (!found || d == 0.0) {
        // the reference leaves such entries uninitialised
        // (preconditioner_jacobi.cpp:39-49); here they act as identity
        atomicAdd(missing, 1u);
        idiag[i] = 1.0;
        idiag32[i] = 1.0f;
        return;
    }
    idiag[i] = 1.0 / d;
    idiag32[i] = v64 ? (float)(1.0 / d) : 1.0f / (float)d;
}

template <class F> int32_t g(F&& f) { return guarded(std::forward<F>(f)); }

// Run a solve; a solve that ends above rtol returns kErrNotConverged with
// its outputs and stats written.
template <class F> int32_t not_converged_status(SolverCG* sol, double rtol, F&& f) {
    const int32_t st = guarded(std::forward<F>(f));
    if (st) return st;
    if (!(sol->last.true_rel <= rtol)) {
        char buf[160];
        std::snprintf(buf, sizeof(buf), "CG stopped at relative residual %.3e above %.3e after %u iterations",
                      sol->last.true_rel, rtol, sol->last.iters);
        set_last_error(buf);
        return kErrNotConverged;
    }
    return kOk;
}
}  // namespace

extern "C" {

d_CSR* d_csr_create(void) {
    try {
        (void)ctx();
        return new d_CSR();
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return nullptr;
    }
}
void d_csr_destroy(d_CSR* A) { delete A; }

h_CSR* h_csr_create(uint32_t n, const uint32_t* row, const uint32_t* col, const float* val) {
    auto* h = new h_CSR();
    h->n = n;
    const uint32_t nnz = row[n];
    h->row.assign(row, row + n + 1);
    h->col.assign(col, col + nnz);
    h->val32.assign(val, val + nnz);
    return h;
}
// FP64 values (extension).
h_CSR* h_csr_create_f64(uint32_t n, const uint32_t* row, const uint32_t* col, const double* val) {
    auto* h = new h_CSR();
    h->n = n;
    const uint32_t nnz = row[n];
    h->row.assign(row, row + n + 1);
    h->col.assign(col, col + nnz);
    h->val64.assign(val, val + nnz);
    h->f64 = true;
    return h;
}
void h_csr_destroy(h_CSR* A) { delete A; }

void d_csr_h2d(d_CSR* d, const h_CSR* h) {
    int32_t st = g([&] {
        if (!d || !h) throw std::invalid_argument("null CSR handle");
        const uint32_t n = h->n, nnz = h->row[n];
        d->A.n = n;
        d->A.nnz = nnz;
        d->A.row.resize(n + 1);
        d->A.col.resize(nnz ? nnz : 1);
        d->val32.resize(nnz ? nnz : 1);
        h2d(d->A.row.p, h->row.data(), n + 1);
        h2d(d->A.col.p, h->col.data(), nnz);
        const auto v32 = h_vec_f32(h);
        h2d(d->val32.p, v32.data(), nnz);
        d->f64 = h->f64;
        d->s32_ok = d->s64_ok = false;
        if (h->f64) {
            d->A.val.resize(nnz ? nnz : 1);
            h2d(d->A.val.p, h->val64.data(), nnz);
        }
    });
    (void)st;
}

LinSys* LinSys_create(void) { return new LinSys(); }
void LinSys_destroy(LinSys* s) { delete s; }

void LinSys_set(LinSys* sys, const d_CSR* A, const float* b) {
    (void)g([&] {
        sys->n = A->A.n;
        sys->A = A;
        std::vector<double> bd(b, b + A->A.n);
        sys->b.resize(sys->n ? sys->n : 1);
        h2d(sys->b.p, bd.data(), sys->n);
    });
}
// FP64 right-hand side (extension).
int32_t LinSys_set_f64(LinSys* sys, const d_CSR* A, const double* b) {
    return g([&] {
        sys->n = A->A.n;
        sys->A = A;
        sys->b.resize(sys->n ? sys->n : 1);
        h2d(sys->b.p, b, sys->n);
    });
}

PrecJacobi_st* PrecJacobi_create(const d_CSR* A) {
    auto* p = new PrecJacobi_st();
    int32_t st = g([&] {
        const uint32_t n = A->A.n;
        p->idiag.resize(n ? n : 1);
        p->idiag32.resize(n ? n : 1);
        DBuf<uint32_t> miss;
        miss.resize(1);
        fill<uint32_t>(miss.p, 1, 0u, ctx().stream);
        hipLaunchKernelGGL(jacobi_kernel, dim3(grid_for(n, 256)), dim3(256), 0, ctx().stream, A->A.row.p, A->A.col.p,
                           A->val32.p, A->f64 ? A->A.val.p : nullptr, n, p->idiag.p, p->idiag32.p, miss.p);
        p->missing = read_scalar(miss.p);
    });
    if (st) {
        delete p;
        return nullptr;
    }
    return p;
}
void PrecJacobi_destroy(PrecJacobi_st* p) { delete p; }
// Rows without a usable diagonal (treated as identity).
uint32_t PrecJacobi_missing_diagonals(const PrecJacobi_st* p) { return p ? p->missing : 0u; }

SolverCG* SolverCG_create(void) {
    try {
        (void)ctx();
        return new SolverCG();
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return nullptr;
    }
}
void SolverCG_destroy(SolverCG* s) { delete s; }

// Reference FP32 semantics (conjugate_gradient.cpp:61-136,138-208): stop
// when sqrt(r.z) (Jacobi) or sqrt(r.r) (plain) <= 1e-6, at most 10001
// loop bodies, and return the reference's post-incremented count.
static uint32_t ref_solve(SolverCG* sol, const LinSys* sys, const PrecJacobi_st* prec, float* x) {
    uint32_t ret = 0;
    (void)g([&] {
        const d_CSR* A = sys->A;
        const uint32_t n = sys->n;
        sol->x.resize(n ? n : 1);
        std::vector<double> xd(x, x + n);
        h2d(sol->x.p, xd.data(), n);
        d_CSR* Am = const_cast<d_CSR*>(A);
        Am->ensure_f64();
        CgOptions o;
        o.abs_tol2 = 1e-12;
        o.maxit = 10001;
        o.crit = prec ? 1u : 0u;
        o.verify = false;
        o.min_one = true;
        cg_solve<float>(Am->A, Am->sell32(), false, prec ? kPrecJacobi : kPrecNone, prec ? prec->idiag32.p : nullptr,
                        nullptr, sys->b.p, sol->x.p, o, sol->work, sol->last, ctx().stream);
        d2h(xd.data(), sol->x.p, n);
        for (uint32_t i = 0; i < n; ++i) x[i] = (float)xd[i];
        ret = sol->last.iters >= 10001 ? 10001u : sol->last.iters - 1;
    });
    return ret;
}

uint32_t SolverCG_jacobi_solve(SolverCG* solver, const LinSys* sys, const PrecJacobi_st* prec, float* x) {
    if (!prec) {
        set_last_error("SolverCG_jacobi_solve: null preconditioner");
        return 0;
    }
    return ref_solve(solver, sys, prec, x);
}

uint32_t SolverCG_solve(SolverCG* solver, const LinSys* sys, float* x) { return ref_solve(solver, sys, nullptr, x); }

// FP64 Jacobi-PCG to a relative true residual ||b-Ax||/||b|| <= rtol
// (extension).  x: initial guess in, solution out.  stats (may be null):
// [iterations, true relative residual, milliseconds, replacements].
int32_t SolverCG_solve_rel_f64(SolverCG* sol, const LinSys* sys, const PrecJacobi_st* prec, double* x, double rtol,
                               uint32_t maxit, double* stats) {
    return not_converged_status(sol, rtol, [&] {
        const d_CSR* A = sys->A;
        const uint32_t n = sys->n;
        const_cast<d_CSR*>(A)->ensure_f64();
        sol->x.resize(n ? n : 1);
        h2d(sol->x.p, x, n);
        CgOptions o;
        o.rtol = rtol;
        o.maxit = maxit;
        o.batch = tuning().batch;
        d_CSR* Am = const_cast<d_CSR*>(A);
        cg_solve<double>(Am->A, Am->sell64(), false, kPrecJacobi, prec->idiag.p, nullptr, sys->b.p, sol->x.p, o,
                         sol->work, sol->last, ctx().stream);
        d2h(x, sol->x.p, n);
        if (stats) {
            stats[0] = sol->last.iters;
            stats[1] = sol->last.true_rel;
            stats[2] = sol->last.ms;
            stats[3] = sol->last.replacements;
        }
    });
}

// Mixed precision (extension): FP32 matrix/vectors/preconditioner, FP64
// dots, FP64 solution and residual replacement.  Same stats layout.
int32_t SolverCG_solve_mixed(SolverCG* sol, const LinSys* sys, const PrecJacobi_st* prec, double* x, double rtol,
                             uint32_t maxit, double* stats) {
    return not_converged_status(sol, rtol, [&] {
        const d_CSR* A = sys->A;
        const uint32_t n = sys->n;
        const_cast<d_CSR*>(A)->ensure_f64();
        sol->x.resize(n ? n : 1);
        h2d(sol->x.p, x, n);
        CgOptions o;
        o.rtol = rtol;
        o.maxit = maxit;
        o.replace_factor = tuning().replace_factor;
        o.batch = tuning().batch;
        d_CSR* Am = const_cast<d_CSR*>(A);
        cg_solve<float>(Am->A, Am->sell32(), false, kPrecJacobi, prec->idiag32.p, nullptr, sys->b.p, sol->x.p, o,
                        sol->work, sol->last, ctx().stream);
        d2h(x, sol->x.p, n);
        if (stats) {
            stats[0] = sol->last.iters;
            stats[1] = sol->last.true_rel;
            stats[2] = sol->last.ms;
            stats[3] = sol->last.replacements;
        }
    });
}

// Process-wide solver knobs: "replace_factor" (mixed precision: replace
// the recursive residual by the FP64 true residual whenever its norm fell
// by this factor, default 0.1), "batch" (CG iterations between host checks,
// default 32), "cg_blocks" (grid cap of the CG kernels), "graph" (replay
// each batch as a captured HIP graph), "xcd_map" (SELL slice mapping),
// "fold_group" (two-level reduction group size, 0 = one level),
// "sell_chunk" (SELL entries per load step, 4 or 8) and "llt_threads"
// (block size of the L L^T path's kernels, 256, 512 or 1024).
int32_t Stream_tune(const char* key, double value) {
    return g([&] {
        const std::string k(key ? key : "");
        if (k == "replace_factor") {
            if (!(value > 0.0 && value < 1.0)) throw std::invalid_argument("replace_factor must be in (0, 1)");
            tuning().replace_factor = value;
        } else if (k == "batch") {
            if (!(value >= 1 && value <= 4096)) throw std::invalid_argument("batch must be in [1, 4096]");
            tuning().batch = (uint32_t)value;
        } else if (k == "cg_blocks") {
            if (!(value >= 64 && value <= 4096)) throw std::invalid_argument("cg_blocks must be in [64, 4096]");
            tuning().cg_blocks = (uint32_t)value;
        } else if (k == "graph") {
            if (!(value == 0.0 || value == 1.0)) throw std::invalid_argument("graph must be 0 or 1");
            tuning().graph = (uint32_t)value;
        } else if (k == "sell_chunk") {
            if (!(value == 4.0 || value == 8.0)) throw std::invalid_argument("sell_chunk must be 4 or 8");
            tuning().sell_chunk = (uint32_t)value;
        } else if (k == "llt_threads") {
            if (!(value == 256.0 || value == 512.0 || value == 1024.0))
                throw std::invalid_argument("llt_threads must be 256, 512 or 1024");
            tuning().llt_threads = (uint32_t)value;
        } else if (k == "fold_group") {
            set_fold_group((uint32_t)value);
        } else if (k == "edge_capacity") {
            if (!(value >= 1.0 && value <= 6.0 && value == (double)(uint32_t)value))
                throw std::invalid_argument("edge_capacity must be an integer in [1, 6]");
            ashape2d_edge_capacity() = (uint32_t)value;
        } else if (k == "xcd_map") {
            if (!(value == 0.0 || value == 1.0)) throw std::invalid_argument("xcd_map must be 0 or 1");
            tuning().xcd_map = (uint32_t)value;
        } else {
            throw std::invalid_argument("unknown tuning key: " + k);
        }
    });
}

// ===========================================================================
// PfemStep2D: LBVH + alpha-shape + P1 assembly + CG on device-resident
// points (no reference counterpart: the reference never connects mesh and
// solve).
// ===========================================================================
struct PfemStep2D {
    Step2D s;
    int bits = 32;
};

PfemStep2D* PfemStep2D_create(void) {
    try {
        (void)ctx();
        return new PfemStep2D();
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return nullptr;
    }
}
void PfemStep2D_destroy(PfemStep2D* h) { delete h; }

// Points (row-major xy [2N]), alpha [N] and nodal source f [N], input
// order; copied once into device memory.
int32_t PfemStep2D_set_points(PfemStep2D* h, uint32_t n, const double* xy, const double* alpha, const double* f) {
    return g([&] { h->s.set_points(n, xy, alpha, f); });
}
int32_t PfemStep2D_set_morton_bits(PfemStep2D* h, int32_t bits) {
    return g([&] {
        if (bits < 1 || bits > 32) throw std::invalid_argument("Morton bits per axis must be in [1, 32]");
        h->bits = bits;
    });
}

// Fourier modes [n_modes][5] for the device source f = sum (ax kx + ay ky)
// cos(kx x + ky y + phi), rows (kx, ky, phi, ax, ay); evaluated now at the
// current points and again after every advection.  n_modes = 0 returns to
// the host-given f of PfemStep2D_set_points.
int32_t PfemStep2D_set_source_modes(PfemStep2D* h, uint32_t n_modes, const double* modes) {
    return g([&] {
        if (n_modes > 32 || (n_modes && !modes)) throw std::invalid_argument("at most 32 source modes");
        h->s.source.n = n_modes;
        std::memcpy(h->s.source.m, modes, sizeof(double) * 5 * n_modes);
        if (n_modes) {
            h->s.update_source(ctx().stream);
            SA_HIP(hipStreamSynchronize(ctx().stream));
        }
    });
}

// Move the points `steps` forward-Euler steps of dt in the divergence-free
// velocity u = (dpsi/dy, -dpsi/dx) of psi = sum a sin(kx x + ky y + phi),
// modes [n_modes][4] rows (kx, ky, phi, a).
int32_t PfemStep2D_advect(PfemStep2D* h, uint32_t n_modes, const double* modes, double dt, uint32_t steps) {
    return g([&] {
        if (n_modes > 32 || (n_modes && !modes)) throw std::invalid_argument("at most 32 velocity modes");
        Modes md;
        md.n = n_modes;
        for (uint32_t k = 0; k < n_modes; ++k) {
            for (int c = 0; c < 4; ++c) md.m[k][c] = modes[4 * k + c];
            md.m[k][4] = 0.0;
        }
        h->s.advect(md, dt, steps);
    });
}

// Nodal velocity [N][2] (input order) of the projection step.
int32_t PfemStep2D_set_velocity(PfemStep2D* h, const double* uv) {
    return g([&] {
        if (!h || !uv) throw std::invalid_argument("PfemStep2D_set_velocity: bad arguments");
        h->s.set_velocity(uv);
    });
}
int32_t PfemStep2D_get_velocity(PfemStep2D* h, double* uv) {
    return g([&] {
        if (!h || !uv || !h->s.has_velocity) throw std::invalid_argument("PfemStep2D_get_velocity: no velocity set");
        d2h(reinterpret_cast<double2*>(uv), h->s.u_in.p, h->s.n);
    });
}

// One fractional (projection) step of the PFEM velocity on the current
// points: mesh + assembly, predictor u* = u + dt (g - nu M^-1 K u),
// pressure K p = -(rho / dt) D(u*) (p = 0 on free-surface nodes) solved by
// `solver` (as PfemStep2D_run) to rtol, correction u = u* - (dt / rho)
// M^-1 G(p); move != 0 then moves the points by dt u.  g [2]: body force.
// stats [11]: PfemStep2D_run's ten, then the ms of the velocity kernels.
int32_t PfemStep2D_projection_step(PfemStep2D* h, int32_t solver, double dt, double rho, double nu, const double* gv,
                                   double rtol, uint32_t maxit, int32_t move, double* stats) {
    return g([&] {
        if (!h || !gv) throw std::invalid_argument("PfemStep2D_projection_step: bad arguments");
        if (solver < 0 || solver > 3) throw std::invalid_argument("unknown solver kind");
        h->s.projection_step(h->bits, solver, dt, rho, nu, gv[0], gv[1], rtol, maxit, move != 0);
        if (stats) {
            stats[0] = h->s.t_mesh;
            stats[1] = h->s.t_asm;
            stats[2] = h->s.t_solve;
            stats[3] = h->s.last.iters;
            stats[4] = h->s.last.true_rel;
            stats[5] = h->s.last.replacements;
            stats[6] = h->s.mesh.n_edges() + h->s.n;
            stats[7] = h->s.mesh.n_elems();
            stats[8] = h->s.t_prep;
            stats[9] = h->s.learned_fallback ? 1.0 : 0.0;
            stats[10] = h->s.t_flow;
        }
        check_error_flag();
        if (h->s.last.true_rel > rtol) throw std::runtime_error("CG did not reach the requested residual");
    });
}

// Fields of the last projection step in Morton order (any pointer may be
// null): u* [N][2], the new velocity [N][2], the lumped mass [N] and the
// discrete divergence D(u_new) [N].
int32_t PfemStep2D_get_flow_fields(PfemStep2D* h, double* us, double* unew, double* mass, double* div) {
    return g([&] { h->s.flow_fields(us, unew, mass, div); });
}

// Current point coordinates [N][2] (input order).
int32_t PfemStep2D_get_points(PfemStep2D* h, double* xy) {
    return g([&] {
        std::vector<Sph4<double>> t(h->s.n);
        d2h(t.data(), h->s.pts.p, h->s.n);
        for (uint32_t i = 0; i < h->s.n; ++i) {
            xy[2 * i] = t[i].c[0];
            xy[2 * i + 1] = t[i].c[1];
        }
    });
}

// Network used by solver 3 (not owned; null detaches).
int32_t PfemStep2D_set_learned(PfemStep2D* h, PrecLearned* net) {
    return g([&] { h->s.learned = net; });
}

// Mesh and assemble without solving (the system is then available to
// PfemStep2D_get_system and PrecLearned_bind_step).  stats [2] (nullable):
// mesh ms, assembly ms.
int32_t PfemStep2D_assemble(PfemStep2D* h, double* stats) {
    return g([&] {
        h->s.mesh_and_assemble(h->bits, -1);
        if (stats) {
            stats[0] = h->s.t_mesh;
            stats[1] = h->s.t_asm;
        }
        const uint32_t fl = read_scalar(ctx().d_flags + kFlagError);
        if (fl) {
            fill<uint32_t>(ctx().d_flags, 1, 0u, ctx().stream);
            throw std::runtime_error("device error flag raised during the step");
        }
    });
}

// Bind the network's training/inference graph to the step's current
// system (device-resident, no copies).
int32_t PrecLearned_bind_step(PrecLearned* net, PfemStep2D* h) {
    return g([&] {
        if (!net || !h || !h->s.A.nnz) throw std::invalid_argument("PrecLearned_bind_step: no assembled system");
        h->s.bind_graph(net->graph, ctx().stream);
        mpnn_check_flags();
    });
}

// One step: mesh + assemble + solve.  solver: 0 = FP64 Jacobi-PCG,
// 1 = mixed FP32/FP64 Jacobi-PCG, 2 = mixed PCG with the FSAI L L^T factor,
// 3 = mixed PCG with the learned L L^T factor.  stats [10]: mesh ms,
// assembly ms, solve ms, iterations, true relative residual, replacements,
// nnz, n_elems, preconditioner set-up ms (matrix copy + factor), 1 when the
// learned factor was rejected and Jacobi used.
int32_t PfemStep2D_run(PfemStep2D* h, int32_t solver, double rtol, uint32_t maxit, double* stats) {
    return g([&] {
        if (solver < 0 || solver > 3) throw std::invalid_argument("unknown solver kind");
        h->s.mesh_and_assemble(h->bits, solver);
        h->s.solve(solver, rtol, maxit);
        if (stats) {
            stats[0] = h->s.t_mesh;
            stats[1] = h->s.t_asm;
            stats[2] = h->s.t_solve;
            stats[3] = h->s.last.iters;
            stats[4] = h->s.last.true_rel;
            stats[5] = h->s.last.replacements;
            stats[6] = read_scalar(h->s.A.row.p + h->s.n);
            stats[7] = h->s.mesh.n_elems();
            stats[8] = h->s.t_prep;
            stats[9] = h->s.learned_fallback ? 1.0 : 0.0;
        }
        const uint32_t fl = read_scalar(ctx().d_flags + kFlagError);
        if (fl) {
            fill<uint32_t>(ctx().d_flags, 1, 0u, ctx().stream);
            throw std::runtime_error("device error flag raised during the step");
        }
        if (h->s.last.true_rel > rtol) throw std::runtime_error("CG did not reach the requested residual");
    });
}

// Average duration in ms of a CG kernel on the system of the last step,
// timed with HIP events on the library stream over `reps` launches: kind
// 0 = FP64 Jacobi direction kernel (d = z + beta d_old, q = A d, d.q), 1 =
// FP32 row-sum Jacobi direction kernel, 2 = FP32 row-sum L L^T direction
// kernel (q = A d on a ready-made d, d.q), 3 = L L^T update kernel (x +=
// alpha d, r' = r - alpha q, w = L^T r'), 4 = L L^T lower apply (z = L w,
// d_new = z + beta d_old); 3 and 4 on the learned factor when a network is
// set, else the FSAI factor.
int32_t PfemStep2D_time_dir_kernel(PfemStep2D* h, int32_t kind, uint32_t reps, double* ms) {
    return g([&] {
        if (kind < 0 || kind > 4) throw std::invalid_argument("time_dir_kernel: kind must be 0 to 4");
        h->s.require_system();
        CgWork w;
        if (kind >= 3) {
            h->s.prepare(h->s.learned ? 3 : 2, ctx().stream);
            *ms = time_llt_update_kernel<float>(h->s.L32, reps, w, ctx().stream, kind == 4);
            return;
        }
        h->s.prepare(kind ? 1 : 0, ctx().stream);
        if (kind)
            *ms = time_dir_kernel<float>(h->s.S32, true, kind == 2, reps, w, ctx().stream);
        else
            *ms = time_dir_kernel<double>(h->s.S64, false, false, reps, w, ctx().stream);
    });
}

// GPU time of the mesh-and-assembly stages on the current points, averaged
// over `reps` runs (HIP events on the library stream between the stages'
// launches).  ms [11]: global box + Morton codes + digit histograms, radix
// sort (all passes), Morton-order gather, hierarchy, refit, candidate
// search, local Delaunay + filter (+ exact redo), element/edge scans,
// compress, P1 assembly kernel, whole mesh + assembly.
int32_t PfemStep2D_time_sections(PfemStep2D* h, uint32_t reps, double* ms) {
    return g([&] {
        if (!h || !ms || !reps) throw std::invalid_argument("PfemStep2D_time_sections: bad arguments");
        static const int a[11] = {kMarkBuild0, kMarkMorton, kMarkSort, kMarkGather, kMarkKarras, kMarkCompute0,
                                  kMarkFill, kMarkDelaunay, kMarkScans, kMarkAsm0, kMarkBuild0};
        static const int b[11] = {kMarkMorton, kMarkSort, kMarkGather, kMarkKarras, kMarkRefit, kMarkFill,
                                  kMarkDelaunay, kMarkScans, kMarkCompress, kMarkAsm, kMarkAsm};
        for (int k = 0; k < 11; ++k) ms[k] = 0.0;
        struct Disarm {
            ~Disarm() { sections_arm(false); }
        } disarm;
        for (uint32_t r = 0; r < reps; ++r) {
            sections_arm(true);
            h->s.mesh_and_assemble(h->bits, -1);
            for (int k = 0; k < 11; ++k) ms[k] += sections_elapsed(a[k], b[k]);
        }
        sections_arm(false);
        for (int k = 0; k < 11; ++k) ms[k] /= reps;
        check_error_flag();
    });
}

// Solution in input order [N].
int32_t PfemStep2D_get_solution(PfemStep2D* h, double* p) {
    return g([&] { h->s.solution_input_order(p); });
}

// Solution in Morton order [N].
int32_t PfemStep2D_get_solution_morton(PfemStep2D* h, double* p) {
    return g([&] {
        h->s.require_solution();
        d2h(p, h->s.x.p, h->s.n);
    });
}

uint32_t PfemStep2D_nnz(PfemStep2D* h) {
    uint32_t r = 0;
    (void)g([&] {
        h->s.require_system();
        r = read_scalar(h->s.A.row.p + h->s.n);
    });
    return r;
}

// Assembled system in Morton order: row [N+1], col/val [nnz], b [N].
int32_t PfemStep2D_get_system(PfemStep2D* h, uint32_t* row, uint32_t* col, double* val, double* b) {
    return g([&] {
        h->s.require_system();
        const uint32_t n = h->s.n;
        const uint32_t nnz = read_scalar(h->s.A.row.p + n);
        if (row) d2h(row, h->s.A.row.p, n + 1);
        if (col) d2h(col, h->s.A.col.p, nnz);
        if (val) d2h(val, h->s.A.val.p, nnz);
        if (b) d2h(b, h->s.b.p, n);
    });
}

// Mesh outputs of the last step (Morton ids): elements [3*nelem],
// node_is_bnd [N], Morton-ordered xy [2N], permutation [N].  Any may be null.
int32_t PfemStep2D_get_mesh(PfemStep2D* h, uint32_t* elem, uint8_t* node_is_bnd, double* xy, uint32_t* perm) {
    return g([&] {
        auto& m = h->s.mesh;
        const uint32_t n = m.n;
        if (elem) {
            const uint32_t ne = m.n_elems();
            if (ne > m.cap_elems) throw std::runtime_error("triangle count exceeds the mesh's capacity");
            d2h(elem, m.elemglob.p, 3ull * ne);
        }
        if (node_is_bnd) d2h(node_is_bnd, m.node_is_bnd.p, n);
        if (perm) d2h(perm, m.lbvh.perm, n);
        if (xy) {
            std::vector<Sph4<double>> v(n);
            d2h(v.data(), m.lbvh.obj_m.p, n);
            for (uint32_t i = 0; i < n; ++i) {
                xy[2 * i] = v[i].c[0];
                xy[2 * i + 1] = v[i].c[1];
            }
        }
    });
}


// Nodal fields of the current points between input order and the Morton
// order of the step's mesh, matrix and solution (permuteForward /
// permuteBackward of the alpha-shape): forward out[i] = in[perm[i]],
// backward out[perm[i]] = in[i]; host arrays of n elements of elem_bytes.
// Needs a meshed step.
int32_t PfemStep2D_permute_forward(PfemStep2D* h, const void* in, void* out, uint32_t elem_bytes) {
    return g([&] {
        if (!h || !h->s.mesh.inited || h->s.mesh.n != h->s.n) throw std::invalid_argument("PfemStep2D: no mesh for the current points");
        permute_host_elements(in, out, h->s.mesh.lbvh.perm, h->s.n, elem_bytes, true);
    });
}
int32_t PfemStep2D_permute_backward(PfemStep2D* h, const void* in, void* out, uint32_t elem_bytes) {
    return g([&] {
        if (!h || !h->s.mesh.inited || h->s.mesh.n != h->s.n) throw std::invalid_argument("PfemStep2D: no mesh for the current points");
        permute_host_elements(in, out, h->s.mesh.lbvh.perm, h->s.n, elem_bytes, false);
    });
}

}  // extern "C"
