// PCG driver interfaces (device-resident scalars, chunked host checks).
#pragma once
#include "sgipc_internal.hpp"

namespace sgipc {

// Device-resident solver state.  Written by kernels, read by the host once
// per chunk of iterations.
struct PcgCtl {
    double rz0;
    double bb;
    double tol;
    double final_rr;
    double final_rz;
    double last_alpha;
    int stop_mode;
    int done;
    int iters;
    int not_spd;
    double diverge;  // > 0: stop (diverged = 1) once ||r||^2 > diverge ||b||^2
    int diverged;
};

template <typename TV>
struct SpmvOp;


// The stop rule evaluated after an x/r update, from the summed r.z and r.r:
// mode 0 |rz| <= tol rz0 (the reference rule), mode 1 ||r||^2 <= tol^2 ||b||^2.
// Short of convergence, rz <= 0 (the preconditioner is not positive definite
// on r) stops with not_spd, and under mode 1 with a divergence bound set an
// r.r above it stops with diverged.  Block 0's thread 0 records the exit.
// Returns whether the run stops.
__device__ __forceinline__ bool pcg_stop_verdict(PcgCtl* ctl, double rz, double rr)
{
    const bool stop = (ctl->stop_mode == 0) ? fabs(rz) <= ctl->tol * ctl->rz0 : rr <= ctl->tol * ctl->tol * ctl->bb;
    const bool bad = !stop && !(rz > 0.0);
    const bool div = !stop && !bad && ctl->stop_mode != 0 && ctl->diverge > 0.0 && !(rr <= ctl->diverge * ctl->bb);
    if ((stop || bad || div) && blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->done = 1;
        ctl->final_rr = rr;
        ctl->final_rz = rz;
        if (bad) ctl->not_spd = 1;
        if (div) ctl->diverged = 1;
    }
    return stop || bad || div;
}

// What the first kernel after the x/r update needs to evaluate the stop rule.
struct PcgStopDesc {
    PcgCtl* ctl = nullptr;
    const double* rz = nullptr;  // partials of the rz used for alpha
    int n_rz = 0;
    const double* rr = nullptr;  // partials of ||r||^2 after the update
    int n_rr = 0;
};

struct PcgOptions {
    double tol = 1e-5;
    int stop_mode = 1;     // 0 = reference |rz| <= tol*rz0, 1 = ||r|| <= tol*||b||
    int max_iter = 1000;
    int check_every = 8;
    // 1: fold the x/r update into the preconditioner's first kernel;
    // 2: also fold the p update into the SpMV (gathers z and p_old, which
    // costs more than the separate update on this part); 0: neither
    int fuse = 1;
    // flexible PCG (preconditioner that may change between iterations):
    // beta = -(z_new . Ap) / (p . Ap) instead of rz_new / rz_old
    bool flexible = false;
    // copies of the residual after every record_stride-th iteration (FP64
    // solver only), for training the learned coarse level
    double* record = nullptr;
    int record_stride = 1;
    int record_cap = 0;
    int* record_count = nullptr;  // host
    // residual replacement: every replace_every iterations the recursive
    // residual becomes b - A_true x, with A_true = true_op (the FP64 matrix
    // when the iteration's SpMV reads FP32 values); p and the Krylov
    // recurrences continue
    const SpmvOp<double>* true_op = nullptr;
    int replace_every = 0;
    // capture the kernels of a full chunk (check_every iterations, from the
    // second chunk on, when the loop's state repeats chunk to chunk) into a
    // HIP graph once per solve and replay it for the later full chunks
    bool graph = false;
    // > 0 (true-residual rule only): stop once ||r|| exceeds sqrt(diverge)
    // ||b|| -- an iteration whose operators cannot resolve the system (the
    // FP32 passes of iterative refinement) ends instead of running on
    double diverge = 0.0;
};

// Inputs of the x/r update (x += alpha p, r -= alpha Ap, alpha = rz/pAp) that
// a preconditioner may perform in its first kernel; rr receives ||r||^2
// partials (rr_count of them).
template <typename TV>
struct FusedXR {
    PcgCtl* ctl = nullptr;
    const double* rz = nullptr;
    int n_rz = 0;
    const double* pap = nullptr;
    int n_pap = 0;
    TV* x = nullptr;
    TV* r = nullptr;
    const TV* p = nullptr;
    const TV* Ap = nullptr;
    double* rr = nullptr;
    int rr_cap = 0;  // capacity of rr
    long n3 = 0;     // length of x, r, p, Ap
};

struct PcgResult {
    int iterations = 0;
    bool converged = false;
    bool not_spd = false;
    bool diverged = false;
    double rel_residual = 0.0;  // recursive residual at exit
    double rz = 0.0;
    double rz0 = 0.0;
};

// Ap = A p (Ap must be zero on entry) + per-workgroup partials of p.Ap; returns
// the partial count.  Kernels return at once when ctl->done is set.
template <typename TV>
struct SpmvOp {
    int nnz = 0;
    int max_partials = 0;  // partials apply_dot may write
    virtual ~SpmvOp() = default;
    virtual int apply_dot(hipStream_t s, const TV* p, TV* Ap, double* pap_part, const PcgCtl* ctl) const = 0;
    // p_new = z + beta p_old (beta = rz_new / rz_old from partials) fused
    // into Ap = A p_new; returns the p.Ap partial count.  Called only when
    // fuses_update_p() is true.
    virtual bool fuses_update_p() const { return false; }
    virtual int apply_dot_update_p(hipStream_t, const TV*, const TV*, TV*, TV*, double*, const PcgCtl*,
                                   const double*, const double*, int) const
    {
        return -1;
    }
    // r = b - A x (residual replacement); false when not offered
    virtual bool residual(hipStream_t, const TV*, const TV*, TV*, const PcgCtl*) const { return false; }
};

// z = M r + per-workgroup partials of r.z; evaluates `stop` first when its ctl
// is set (and then writes nothing).  Returns the partial count, which must not
// depend on the call.
template <typename TV>
struct Preconditioner {
    virtual ~Preconditioner() = default;
    virtual int apply(hipStream_t s, const TV* r, TV* z, double* rz_part, const PcgStopDesc& stop) = 0;
    // x/r update fused into the first kernel, z = M r with r.z partials; the
    // stop rule (previous rz, the new rr) is evaluated by a later kernel of
    // the apply.  Returns the rz partial count (and *n_rr), or -1 when the
    // preconditioner has no fused form.
    virtual int apply_update_xr(hipStream_t, const FusedXR<TV>&, TV*, double*, int*) { return -1; }
    // upper bound of the partial count apply() returns
    virtual int max_partials() const { return 1024; }
    // upper bound of the rr partial count apply_update_xr() writes
    virtual int max_rr_partials() const { return 0; }
    // flexible PCG: where apply() writes the z.Ap partials (same count as
    // the r.z partials); false when the preconditioner cannot
    virtual bool set_flex(const TV*, double*) { return false; }
};

template <typename TV>
class PcgSolver {
  public:
    ~PcgSolver();
    PcgResult solve(hipStream_t s, const SpmvOp<TV>& A, Preconditioner<TV>* M, const TV* b, TV* x,
                    long n3, const PcgOptions& opt);
    void reserve(long n3, int nnz, int max_partials);

  private:
    DevBuf<TV> x_, r_, z_, p_, p2_, Ap_;
    DevBuf<double> part_rz_[2], part_pap_, part_rr_, part_bb_, part_zap_;
    DevBuf<PcgCtl> ctl_;
    PcgCtl* host_ctl_ = nullptr;
    hipEvent_t ev_[2] = {nullptr, nullptr};
};

int vec_grid(long n);
void launch_axpy_f64(hipStream_t s, double a, const double* x, double* y, long n);
// stream n doubles through the caches (read only; sink is never written in practice)
void launch_stream_read(hipStream_t s, const double* buf, long n, double* sink);
template <typename TA, typename TB>
void launch_convert(hipStream_t s, const TA* a, TB* b, long n);
template <typename T>
void launch_fill(hipStream_t s, T* a, T v, long n);
template <typename TA, typename TB>
int launch_dot_partials(hipStream_t s, const TA* a, const TB* b, long n, double* part);
void launch_residual(hipStream_t s, const double* vals, const int* rows, const int* cols, int nnz,
                     int n_rows, const double* x, const double* b, double* r, double* rr_part,
                     int* n_rr);

// SpMV kernels that honour the PCG done flag.
void launch_spmv_pcg_f64(hipStream_t s, const double* vals, const int* rows, const int* cols,
                         int nnz, const double* p, double* Ap, double* pap_part, const PcgCtl* ctl,
                         int* n_part);
void launch_spmv_pcg_f32(hipStream_t s, const float* vals, const int* rows, const int* cols,
                         int nnz, const float* p, float* Ap, double* pap_part, const PcgCtl* ctl,
                         int* n_part);

}  // namespace sgipc
