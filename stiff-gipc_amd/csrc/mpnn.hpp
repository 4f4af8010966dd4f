// Learned coarse level: a message-passing network over the MAS level-1
// aggregate graph, run with BF16 MFMA, that adds a coarse correction to the
// MAS output inside flexible PCG.  No counterpart in the reference.
//
// Model (C = 64 channels, L layers, one row per node component, 3 rows per
// node):
//   Ahat = S A_1 S, A_1 = P^T A P the level-1 Galerkin operator, S = diag(s),
//   s_k = (tr A_1[k,k] / 3)^-1/2.
//   X = [ghat, zhat, rho] per row: ghat = S P^T r, zhat = S^-1 (coarse part of
//   the MAS output at level 1), rho = ghat - Ahat zhat.
//   V0 = X W_in;  V <- V + act([V, Ahat V] W_l);  chat = V w_out, act = relu or
//   the identity.
//   Correction c = S chat, prolonged and added to the MAS output.
//   Loss (0.5 chat^T Ahat chat - chat^T rho) / rows per sample: the A_1
//   energy of the coarse part of the preconditioned vector, measured from
//   the MAS coarse part; its minimiser makes the coarse part the exact
//   level-1 solve A_1^-1 P^T r.
#pragma once
#include "sgipc_internal.hpp"

#include <cstdint>

namespace sgipc {

constexpr int kMpnnC = 64;
constexpr int kMpnnMaxLayers = 4;
constexpr int kMpnnTile = 32;  // nodes per workgroup in the layer kernels (96 rows)

// Level-1 graph: CSR over nodes with both directions and the diagonal,
// columns sorted, blocks Jacobi-scaled (row-major 3x3, float).
struct CoarseGraph {
    int n = 0;
    int nnz = 0;
    DevBuf<int> row_ptr, col;
    DevBuf<float> blk;
    DevBuf<float> s;
    // from the upper BCOO of A_1 (FP64 column-major blocks, sorted unique,
    // indices in [0, n)); deterministic
    void build(hipStream_t st, const double* vals, const int* rows, const int* cols, int nb, int n_nodes);
    void copy_from(hipStream_t st, const CoarseGraph& o);

  private:
    DevBuf<uint64_t> keys_, skeys_;
    DevBuf<int> idx_, sidx_, cnt_;
    DevBuf<unsigned char> temp_;
};

// Network parameters and the BF16 copies the MFMA kernels read.
class Mpnn {
  public:
    // linear != 0: identity activation (the network is linear in its inputs)
    Mpnn(Context* ctx, int layers, int linear);
    int linear() const { return linear_; }
    int layers() const { return L; }
    static int n_params(int layers) { return 3 * kMpnnC + layers * 2 * kMpnnC * kMpnnC + kMpnnC; }
    int n_params() const { return n_params(L); }
    void set_params(const float* host);
    void get_params(float* host) const;
    void init_random(uint64_t seed);

    // chat for n_samples stacked samples (X: [S][3n][3] float, chat: [S][3n]);
    // with scaled != nullptr (S == 1) also writes s .* chat there.
    void forward(hipStream_t st, const CoarseGraph& g, const float* X, int n_samples, float* chat,
                 float* scaled, const int* done = nullptr);
    // mean loss over the samples and its gradient (device, n_params floats)
    double loss_grad(hipStream_t st, const CoarseGraph& g, const float* X, const float* q, int n_samples,
                     float* grad);
    // one Adam step with the gradient in grad_ (after loss_grad)
    void adam_step(hipStream_t st, float lr, float b1, float b2, float eps);
    const float* grad() const { return grad_.get(); }
    float* grad_buf() { return grad_.get(); }
    void reset_optimizer();

    Context* ctx;

  private:
    void sync_bf16(hipStream_t st);
    void reserve_train(long rows_total, int tiles_total);
    int L;
    int linear_ = 0;
    int step_ = 0;
    DevBuf<float> w_;       // flat parameters (W_in | W_0..W_{L-1} | w_out)
    DevBuf<uint16_t> wt_;   // per layer W^T as bf16, [C][2C]
    DevBuf<uint16_t> wn_;   // per layer W as bf16, [2C][C]
    DevBuf<float> m_, v_, grad_;
    // activations
    DevBuf<float> V_[kMpnnMaxLayers + 1];
    DevBuf<uint16_t> U_[kMpnnMaxLayers];
    DevBuf<float> Y_[kMpnnMaxLayers];
    DevBuf<float> G_, G2_, dU_, AV_;
    DevBuf<uint16_t> dY_;
    DevBuf<float> chat_, dchat_, part_;
    DevBuf<double> lpart_;
    long rows_cap_ = 0;
    bool dirty_ = true;  // bf16 copies stale
};

}  // namespace sgipc
