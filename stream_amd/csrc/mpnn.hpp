// Learned sparse-approximate-inverse factor: a message-passing network that
// emits the values of an upper-triangular Ghat on the pattern of the
// pressure matrix A, giving M^-1 = L L^T with L = D^-1/2 Ghat^T (SPD by
// construction: Ghat_kk = softplus(.) > 0).  No counterpart in the
// reference, whose only preconditioner is Jacobi
// (include/numerics/preconditioner.hpp:60-72).
//
// Network (hidden 64, 3 message-passing layers, weights in BF16 on MFMA,
// FP32 accumulation, FP32 master weights + Adam):
//   F_i  = [(x_i-xmin)/ext, (y_i-ymin)/ext, (z_i-zmin)/ext, bnd_i, A_ii/8]
//          (ext = largest box extent; the z feature is 0 for 2D systems)
//   e_ij = [Ahat_ij, |x_i-x_j|/lbar, sign(j-i)]
//   H0 = relu(F W_enc + b_enc)
//   layer: [P|Q] = H W_pq;  M_i = mean_j relu(P_i + Q_j + e_ij W_e + b_m)
//          U = [H|M] W_u + b_u;  H <- H + relu(U)
//   head:  [R|S|Dg] = H W_dec
//          Ghat_ki (i > k) = relu(R_k + S_i + e_ki W_de + b_de) . w_o  (0 where A_ki = 0)
//          Ghat_kk = 1 (Dirichlet) or softplus(relu(Dg_k + b_dd) . w_d + log(e-1))
// Loss: Kaporin's condition number of Ghat Ahat Ghat^T per node,
//   log(tr(Ghat Ahat Ghat^T)/n) - (2/n) sum log Ghat_kk,
// whose minimiser over the pattern is the classical FSAI factor.
//
// Node-level dense layers run as MFMA GEMMs (v_mfma_f32_16x16x32_bf16,
// 64 rows per workgroup, weights staged in LDS); edge work (messages,
// per-edge head, loss) is one wavefront per node with the 64 features on
// the 64 lanes.  Backward kernels recompute edge pre-activations instead of
// storing per-edge tensors, and every reduction over nodes goes through
// per-block partials folded in a fixed order, so training is deterministic.
#pragma once
#include "numerics.hpp"

namespace stream_amd {

constexpr uint32_t kHid = 64;
constexpr uint32_t kLayers = 3;
constexpr uint32_t kNodeFeats = 5;
constexpr uint32_t kEdgeFeats = 3;
constexpr uint32_t kMaxUpper = 31;  // upper entries per row handled by the loss kernel
constexpr int kFlagLearned = 1;      // d_flags word: non-finite learned factor

// Flat parameter layout (FP32 master copy), in this order:
// enc_w[5][64] enc_b[64] | per layer: pq_w[64][128] e_w[3][64] m_b[64]
// u_w[128][64] u_b[64] | dec_w[64][192] de_w[3][64] de_b[64] o_w[64]
// dd_b[64] d_w[64].  Matrices are [in][out] row-major.
struct ParamOffsets {
    uint32_t enc_w, enc_b;
    uint32_t pq_w[kLayers], e_w[kLayers], m_b[kLayers], u_w[kLayers], u_b[kLayers];
    uint32_t dec_w, de_w, de_b, o_w, dd_b, d_w;
    uint32_t total;
};

constexpr ParamOffsets param_offsets() {
    ParamOffsets p{};
    uint32_t o = 0;
    p.enc_w = o; o += kNodeFeats * kHid;
    p.enc_b = o; o += kHid;
    for (uint32_t l = 0; l < kLayers; ++l) {
        p.pq_w[l] = o; o += kHid * 2 * kHid;
        p.e_w[l] = o; o += kEdgeFeats * kHid;
        p.m_b[l] = o; o += kHid;
        p.u_w[l] = o; o += 2 * kHid * kHid;
        p.u_b[l] = o; o += kHid;
    }
    p.dec_w = o; o += kHid * 3 * kHid;
    p.de_w = o; o += kEdgeFeats * kHid;
    p.de_b = o; o += kHid;
    p.o_w = o; o += kHid;
    p.dd_b = o; o += kHid;
    p.d_w = o; o += kHid;
    p.total = o;
    return p;
}
constexpr ParamOffsets kParams = param_offsets();

// Derived per-matrix data.  row/col/val/xy/bnd point to device arrays owned
// elsewhere (the step driver or the network's own copies).
struct MpnnGraph {
    uint32_t n = 0;
    uint32_t nnz = 0;
    const uint32_t* row = nullptr;
    const uint32_t* col = nullptr;
    const double* val = nullptr;
    const double* xy = nullptr;  // point i at xy[i * xy_stride + 0 .. dim-1]
    uint32_t xy_stride = 2;
    uint32_t dim = 2;            // 2 or 3
    const uint8_t* bnd = nullptr;
    DBuf<double> sd;          // [n] 1/sqrt(A_ii) (1 when A_ii <= 0)
    DBuf<uint32_t> dpos;      // [n] CSR position of the diagonal (~0u if absent)
    DBuf<float> deg;          // [n] off-diagonal count
    DBuf<float> nfeat;        // [n][kNodeFeats]
    DBuf<float2> efeat;       // [nnz] (Ahat_ij, len_ij/lbar), rounded to bf16
    DBuf<uint32_t> tr;        // [nnz] position of (j, i) for entry (i, j)
    bool has_transpose = false;  // tr built (the backward pass needs it)
    DBuf<double> part;        // reduction partials
    DBuf<double> scal;        // [8] xmin, xmax, ymin, ymax, zmin, zmax, len sum, count
    // owned copies for host-provided systems
    DBuf<uint32_t> own_row, own_col;
    DBuf<double> own_val, own_xy;
    DBuf<uint8_t> own_bnd;
};

// Features, scaling and (transpose = true) the transpose map of the bound
// system; inference does not need the map.
void mpnn_prepare(MpnnGraph& g, hipStream_t s, bool transpose = true);

struct MpnnNet {
    DBuf<float> w, m, v, grad;   // [P] master weights, Adam moments, gradient
    DBuf<uint16_t> wb;           // [P] BF16 copy ([in][out])
    DBuf<uint16_t> wt;           // BF16 [out][in] copies of pq_w, u_w, dec_w
    uint32_t adam_t = 0;
    bool weights_set = false;
    // activations (row-major, FP32)
    DBuf<float> H[kLayers + 1], PQ[kLayers], M[kLayers], U[kLayers], RSD;
    DBuf<float> gval;            // [nnz] Ghat per CSR entry (upper entries and diagonal)
    DBuf<float> od;              // [n] diagonal head pre-softplus
    DBuf<double> y;              // [nnz] (Ghat Ahat) at the pattern of Ghat
    DBuf<float> dRSD, dHa, dHb, dX, dPQ, dU;
    DBuf<float> partials;
    DBuf<double> lpart;          // loss partials
    DBuf<double> lsc;            // [4] T, sum log Ghat_kk, loss, n
    uint32_t cap_n = 0, cap_nnz = 0;

    void set_params(const float* host, hipStream_t s);
    void get_params(float* host);
    void reserve(uint32_t n, uint32_t nnz);
    // keep_pre: store the layers' pre-activations for a backward pass
    // (inference passes false and writes none)
    void forward(const MpnnGraph& g, hipStream_t s, bool keep_pre = true);
    void loss(const MpnnGraph& g, hipStream_t s);      // after forward: lsc, y
    void backward(const MpnnGraph& g, hipStream_t s);  // after loss: grad
    void adam(float lr, float b1, float b2, float eps, hipStream_t s);
    // Write L = D^-1/2 Ghat^T into the factor (structure must match A);
    // raises kFlagLearned on a non-finite or non-positive value.
    void pack(const MpnnGraph& g, LowerFactor<float>& L, hipStream_t s);
};

// Throw (and clear) when a device error flag is set.
void mpnn_check_flags();

}  // namespace stream_amd

// C handle of the learned preconditioner (stream_learned.h).
struct PrecLearned {
    stream_amd::MpnnNet net;
    stream_amd::MpnnGraph graph;
    float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
};
