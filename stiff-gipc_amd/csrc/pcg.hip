// Preconditioned conjugate gradient for gfx950 (PCGSolver::pcg,
// StiffGIPC/linear_system/solver/pcg_solver.cu:191-332).
//
// MI355X mapping.  The reference does two blocking device->host reads per
// iteration (the two dot products) and brackets every phase with an event
// sync.  Here alpha and beta never leave the device: every kernel that needs a
// scalar re-reduces the previous kernel's per-workgroup partials in a fixed
// order (deterministic, no atomics, visibility guaranteed by the kernel
// boundary).  The stop test is evaluated on the device by the first kernel
// after the x/r update, which raises a `done` flag; every later kernel of the
// iteration (and of the iterations already queued behind it) returns at once.
// The host enqueues chunks of iterations and looks at the flag once per chunk,
// while the next chunk is already queued, so the host round trip never stalls
// the GPU.  p.Ap is produced by the SpMV kernel itself and r.z by the
// preconditioner's prolongation kernel, so each iteration touches p, Ap, r, x,
// z once.
#include "pcg.hpp"
#include "wave_ops.hpp"

#include <chrono>
#include <cmath>

namespace sgipc {

namespace {

constexpr int kVecThreads = 256;

__device__ __forceinline__ double sum_partials(const double* __restrict__ a, int n, double* smem)
{
    return block_sum_array(a, n, smem);
}

// Stop rule after the x/r update of iteration k (ctl->done set by block 0).
// mode 0: |rz_k| <= tol * rz0 (the reference rule, evaluated with the rz used
// to form alpha); mode 1: ||r||^2 <= tol^2 ||b||^2.  rz_k <= 0 before either
// holds stops the run with not_spd set.
__device__ __forceinline__ bool eval_stop(const PcgStopDesc& sd, double* smem)
{
    __shared__ int s_done;
    if (sd.ctl == nullptr) return false;
    // one read per workgroup: block 0 of this kernel may raise the flag
    if (threadIdx.x == 0) s_done = sd.ctl->done;
    __syncthreads();
    if (s_done) return true;
    double rz = sum_partials(sd.rz, sd.n_rz, smem);
    double rr = sum_partials(sd.rr, sd.n_rr, smem);
    return pcg_stop_verdict(sd.ctl, rz, rr);
}

// x += alpha p; r -= alpha Ap; rr partials.  alpha = rz / pAp.
template <typename TV>
__global__ __launch_bounds__(kVecThreads) void pcg_update_xr(
    PcgCtl* __restrict__ ctl, const double* __restrict__ rz_part, int n_rz,
    const double* __restrict__ pap_part, int n_pap, TV* __restrict__ x, TV* __restrict__ r,
    const TV* __restrict__ p, const TV* __restrict__ Ap, long n3, double* __restrict__ rr_part)
{
    __shared__ double smem[16];
    if (ctl->done) return;
    double rz = sum_partials(rz_part, n_rz, smem);
    double pap = sum_partials(pap_part, n_pap, smem);
    double alpha = rz / pap;
    double acc = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long)gridDim.x * blockDim.x) {
        double ri = (double)r[i] - alpha * (double)Ap[i];
        x[i] = (TV)((double)x[i] + alpha * (double)p[i]);
        r[i] = (TV)ri;
        acc += ri * ri;
    }
    double s = block_sum(acc, smem);
    if (threadIdx.x == 0) {
        rr_part[blockIdx.x] = s;
        if (blockIdx.x == 0) {
            ctl->iters += 1;
            ctl->last_alpha = alpha;
            if (!(pap > 0.0)) ctl->not_spd = 1;
        }
    }
}

// p = z + beta p.  beta = rz_new / rz_old.
template <typename TV>
__global__ __launch_bounds__(kVecThreads) void pcg_update_p(
    const PcgCtl* __restrict__ ctl, const double* __restrict__ rz_new_part,
    const double* __restrict__ rz_old_part, int n_rz, TV* __restrict__ p, const TV* __restrict__ z,
    TV* __restrict__ Ap, long n3)
{
    __shared__ double smem[16];
    if (ctl->done) return;
    // the first kPer grid strides are loaded before the partial sums
    constexpr int kPer = 4;
    const long i0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    double zv[kPer], pv[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const long i = i0 + j * stride;
        zv[j] = i < n3 ? (double)z[i] : 0.0;
        pv[j] = i < n3 ? (double)p[i] : 0.0;
    }
    double rzn = sum_partials(rz_new_part, n_rz, smem);
    double rzo = sum_partials(rz_old_part, n_rz, smem);
    double beta = rzn / rzo;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const long i = i0 + j * stride;
        if (i < n3) p[i] = (TV)(zv[j] + beta * pv[j]);
    }
    for (long i = i0 + kPer * stride; i < n3; i += stride) {
        p[i] = (TV)((double)z[i] + beta * (double)p[i]);
    }
}

// Flexible p update: beta = -(z.Ap) / (p.Ap), both from partials.
template <typename TV>
__global__ __launch_bounds__(kVecThreads) void pcg_update_p_flex(
    const PcgCtl* __restrict__ ctl, const double* __restrict__ zap_part, int n_zap,
    const double* __restrict__ pap_part, int n_pap, TV* __restrict__ p, const TV* __restrict__ z, long n3)
{
    __shared__ double smem[16];
    if (ctl->done) return;
    const double zap = sum_partials(zap_part, n_zap, smem);
    const double pap = sum_partials(pap_part, n_pap, smem);
    const double beta = -zap / pap;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long)gridDim.x * blockDim.x) {
        p[i] = (TV)((double)z[i] + beta * (double)p[i]);
    }
}

// Identity preconditioner z = r with r.z partials and the stop test.
template <typename TV>
__global__ __launch_bounds__(kVecThreads) void pcg_identity_apply(
    PcgStopDesc sd, const TV* __restrict__ r, TV* __restrict__ z, long n3, double* __restrict__ rz_part)
{
    __shared__ double smem[16];
    if (eval_stop(sd, smem)) return;
    double acc = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long)gridDim.x * blockDim.x) {
        TV v = r[i];
        z[i] = v;
        acc += (double)v * (double)v;
    }
    double s = block_sum(acc, smem);
    if (threadIdx.x == 0) rz_part[blockIdx.x] = s;
}

// Generic dot partials (used at start-up).
template <typename TA, typename TB>
__global__ __launch_bounds__(kVecThreads) void dot_partials_kernel(const TA* __restrict__ a,
                                                                   const TB* __restrict__ b, long n,
                                                                   double* __restrict__ part)
{
    __shared__ double smem[16];
    double acc = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        acc += (double)a[i] * (double)b[i];
    double s = block_sum(acc, smem);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void pcg_init_ctl(PcgCtl* ctl, const double* rz_part, int n_rz, const double* bb_part,
                             int n_bb, double tol, int stop_mode, double diverge)
{
    __shared__ double smem[16];
    double rz = sum_partials(rz_part, n_rz, smem);
    double bb = sum_partials(bb_part, n_bb, smem);
    if (threadIdx.x == 0) {
        ctl->rz0 = rz;
        ctl->bb = bb;
        ctl->tol = tol;
        ctl->stop_mode = stop_mode;
        ctl->done = (bb == 0.0) ? 1 : 0;
        ctl->iters = 0;
        ctl->not_spd = 0;
        ctl->diverge = diverge;
        ctl->diverged = 0;
        ctl->final_rr = bb;
        ctl->final_rz = rz;
        ctl->last_alpha = 0.0;
    }
}

template <typename TA, typename TB>
__global__ void copy_convert_kernel(const TA* __restrict__ a, TB* __restrict__ b, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        b[i] = (TB)a[i];
}

template <typename T>
__global__ void fill_kernel(T* __restrict__ a, T v, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] = v;
}

// r = b - A x in FP64 (outer residual of the refinement loop), with rr partials.
__global__ __launch_bounds__(kVecThreads) void residual_finish_kernel(const double* __restrict__ b,
                                                                      double* __restrict__ r, long n3,
                                                                      double* __restrict__ rr_part)
{
    __shared__ double smem[16];
    double acc = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long)gridDim.x * blockDim.x) {
        double v = b[i] - r[i];
        r[i] = v;
        acc += v * v;
    }
    double s = block_sum(acc, smem);
    if (threadIdx.x == 0) rr_part[blockIdx.x] = s;
}

__global__ void axpy_kernel(double a, const double* __restrict__ x, double* __restrict__ y, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] += a * x[i];
}

// Reads n doubles (16 bytes per lane per step) and writes nothing unless the
// sum hits an impossible value: streams a buffer through the caches without
// leaving dirty lines behind.
__global__ __launch_bounds__(256) void stream_read_kernel(const double2* __restrict__ v, long n2,
                                                          double* __restrict__ sink)
{
    double acc = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (long)gridDim.x * blockDim.x) {
        const double2 t = v[i];
        acc += t.x + t.y;
    }
    if (acc == -1.2345e300) sink[0] = acc;
}

}  // namespace

void launch_stream_read(hipStream_t s, const double* buf, long n, double* sink)
{
    if (n < 2) return;
    stream_read_kernel<<<4096, 256, 0, s>>>(reinterpret_cast<const double2*>(buf), n / 2, sink);
    SGIPC_LAUNCH_CHECK();
}

void launch_axpy_f64(hipStream_t s, double a, const double* x, double* y, long n)
{
    if (n <= 0) return;
    axpy_kernel<<<vec_grid(n), kVecThreads, 0, s>>>(a, x, y, n);
    SGIPC_LAUNCH_CHECK();
}

int vec_grid(long n)
{
    long g = (n + kVecThreads - 1) / kVecThreads;
    if (g > 1024) g = 1024;
    if (g < 1) g = 1;
    return (int)g;
}

template <typename TA, typename TB>
void launch_convert(hipStream_t s, const TA* a, TB* b, long n)
{
    if (n <= 0) return;
    copy_convert_kernel<TA, TB><<<vec_grid(n), kVecThreads, 0, s>>>(a, b, n);
    SGIPC_LAUNCH_CHECK();
}
template void launch_convert<double, float>(hipStream_t, const double*, float*, long);
template void launch_convert<float, double>(hipStream_t, const float*, double*, long);
template void launch_convert<double, double>(hipStream_t, const double*, double*, long);

template <typename T>
void launch_fill(hipStream_t s, T* a, T v, long n)
{
    if (n <= 0) return;
    fill_kernel<T><<<vec_grid(n), kVecThreads, 0, s>>>(a, v, n);
    SGIPC_LAUNCH_CHECK();
}
template void launch_fill<double>(hipStream_t, double*, double, long);
template void launch_fill<float>(hipStream_t, float*, float, long);
template void launch_fill<int>(hipStream_t, int*, int, long);

template <typename TA, typename TB>
int launch_dot_partials(hipStream_t s, const TA* a, const TB* b, long n, double* part)
{
    int g = vec_grid(n);
    dot_partials_kernel<TA, TB><<<g, kVecThreads, 0, s>>>(a, b, n, part);
    SGIPC_LAUNCH_CHECK();
    return g;
}
template int launch_dot_partials<double, double>(hipStream_t, const double*, const double*, long, double*);
template int launch_dot_partials<float, float>(hipStream_t, const float*, const float*, long, double*);

void launch_residual(hipStream_t s, const double* vals, const int* rows, const int* cols, int nnz,
                     int n_rows, const double* x, const double* b, double* r, double* rr_part,
                     int* n_rr)
{
    long n3 = 3L * n_rows;
    launch_fill<double>(s, r, 0.0, n3);
    launch_spmv_f64(s, 1.0, vals, rows, cols, nnz, x, 1.0, r, n_rows, nullptr, nullptr);
    int g = vec_grid(n3);
    residual_finish_kernel<<<g, kVecThreads, 0, s>>>(b, r, n3, rr_part);
    SGIPC_LAUNCH_CHECK();
    *n_rr = g;
}

// ------------------------------------------------------------------ solver
template <typename TV>
void PcgSolver<TV>::reserve(long n3, int nnz, int max_partials)
{
    x_.reserve(n3);
    r_.reserve(n3);
    z_.reserve(n3);
    p_.reserve(n3);
    p2_.reserve(n3);
    Ap_.reserve(n3);
    int g = vec_grid(n3);
    int gs = max_partials > 0 ? max_partials : spmv_partials_needed(nnz > 0 ? nnz : 1);
    int gmax = g > gs ? g : gs;
    if (gmax < 1024) gmax = 1024;
    part_rz_[0].reserve(gmax);
    part_rz_[1].reserve(gmax);
    part_pap_.reserve(gmax);
    part_rr_.reserve(gmax);
    part_bb_.reserve(gmax);
    part_zap_.reserve(gmax);
    ctl_.reserve(1);
    if (!host_ctl_) SGIPC_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_ctl_), 2 * sizeof(PcgCtl)));
}

template <typename TV>
PcgSolver<TV>::~PcgSolver()
{
    if (host_ctl_) (void)hipHostFree(host_ctl_);
    if (ev_[0]) (void)hipEventDestroy(ev_[0]);
    if (ev_[1]) (void)hipEventDestroy(ev_[1]);
}

template <typename TV>
PcgResult PcgSolver<TV>::solve(hipStream_t s, const SpmvOp<TV>& A, Preconditioner<TV>* M,
                               const TV* b, TV* x_out, long n3, const PcgOptions& opt)
{
    PcgResult res{};
    {
        const int pm = M ? M->max_partials() : 0;
        reserve(n3, A.nnz, A.max_partials > pm ? A.max_partials : pm);
        if (M && M->max_rr_partials() > 0) part_rr_.reserve((size_t)M->max_rr_partials());
    }
    if (!ev_[0]) {
        SGIPC_HIP(hipEventCreateWithFlags(&ev_[0], hipEventDisableTiming));
        SGIPC_HIP(hipEventCreateWithFlags(&ev_[1], hipEventDisableTiming));
    }
    TV* x = x_out;
    TV* r = r_.get();
    TV* z = z_.get();
    TV* pbuf[2] = {p_.get(), p2_.get()};
    int pb = 0;  // pbuf[pb] is the current search direction
    TV* Ap = Ap_.get();
    PcgCtl* ctl = ctl_.get();

    launch_fill<TV>(s, x, TV(0), n3);
    SGIPC_HIP(hipMemcpyAsync(r, b, sizeof(TV) * n3, hipMemcpyDeviceToDevice, s));
    int n_bb = launch_dot_partials<TV, TV>(s, b, b, n3, part_bb_.get());
    // flexible PCG: z.Ap partials from the preconditioner's last kernel,
    // switched on before its first apply so every apply emits the same
    // partial layout (Ap zero until the first SpMV); switched off on every
    // way out
    launch_fill<TV>(s, Ap, TV(0), n3);
    const bool flex = opt.flexible && M != nullptr && M->set_flex(Ap, part_zap_.get());
    struct FlexOff {
        Preconditioner<TV>* m;
        bool on;
        ~FlexOff()
        {
            if (on) m->set_flex(nullptr, nullptr);
        }
    } flex_off{M, flex};
    // z = M r (no stop test), rz0 partials
    int cur = 0;
    PcgStopDesc none{};
    int n_rz = M ? M->apply(s, r, z, part_rz_[cur].get(), none) : 0;
    if (!M) {
        int g = vec_grid(n3);
        pcg_identity_apply<TV><<<g, kVecThreads, 0, s>>>(none, r, z, n3, part_rz_[cur].get());
        SGIPC_LAUNCH_CHECK();
        n_rz = g;
    }
    pcg_init_ctl<<<1, kVecThreads, 0, s>>>(ctl, part_rz_[cur].get(), n_rz, part_bb_.get(), n_bb,
                                           opt.tol, opt.stop_mode, opt.diverge);
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(hipMemcpyAsync(pbuf[pb], z, sizeof(TV) * n3, hipMemcpyDeviceToDevice, s));

    // Fused iteration (when both operators offer it):
    //   SpMV:  p_k = z + beta p_{k-1} gathered on the fly, Ap = A p_k, p.Ap
    //   MAS:   x += alpha p_k, r -= alpha Ap, ||r||^2 in the restriction;
    //          stop rule in the first coarse restriction; z = M r, r.z
    // i.e. the p update of iteration k-1 is deferred into the SpMV of k.
    const bool fuse_p = !flex && opt.fuse >= 2 && A.fuses_update_p();
    bool fuse_xr = opt.fuse >= 1 && M != nullptr;  // cleared when the preconditioner declines
    const bool replace = sizeof(TV) == 8 && opt.true_op != nullptr && opt.replace_every > 0;
    bool p_pending = false;                   // p update of the previous iteration not yet applied
    const int gv = vec_grid(n3);
    const int chunk = opt.check_every > 0 ? opt.check_every : 8;
    int issued = 0;  // iterations enqueued
    int pending = -1;
    int round = 0;
    const int max_it = opt.max_iter > 1 ? opt.max_iter - 1 : 1;  // the reference loops k=1..max_iter-1
    bool finished = false;
    // HIP graph of one full chunk: from the second chunk on the loop's state
    // (cur, pb, p_pending, fuse_xr, the replacement positions) is the same at
    // every chunk boundary when the chunk length is even and a multiple of the
    // replacement period, so one capture serves every later full chunk
    bool use_graph = opt.graph && !flex && opt.record == nullptr && (chunk % 2 == 0) &&
                     (!replace || chunk % opt.replace_every == 0);
    hipGraphExec_t gexec = nullptr;
    int n_chunks = 0;
    while (!finished) {
        int todo = max_it - issued;
        if (todo > chunk) todo = chunk;
        const bool full = todo == chunk;
        if (todo > 0 && use_graph && full && gexec != nullptr) {
            SGIPC_HIP(hipGraphLaunch(gexec, s));
            issued += todo;
            for (int it = 0; it < todo; ++it) cur ^= 1;
            if (fuse_p) pb ^= (todo & 1);
        } else if (todo > 0) {
            const bool capture = use_graph && full && n_chunks >= 1;
            const int cur0 = cur, pb0 = pb;
            const bool pp0 = p_pending, fx0 = fuse_xr;
            if (capture) SGIPC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            // an error while capturing must close the capture before the
            // stream is used again
            struct CaptureGuard {
                hipStream_t s;
                bool open;
                ~CaptureGuard()
                {
                    if (open) {
                        hipGraph_t g = nullptr;
                        if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
                    }
                }
            } guard{s, capture};
            for (int it = 0; it < todo; ++it) {
                const int nxt = cur ^ 1;
                int n_pap;
                if (p_pending) {
                    // beta = rz[cur] / rz[nxt]: rz[nxt] still holds the previous rz
                    n_pap = A.apply_dot_update_p(s, z, pbuf[pb], pbuf[pb ^ 1], Ap, part_pap_.get(), ctl,
                                                 part_rz_[cur].get(), part_rz_[nxt].get(), n_rz);
                    pb ^= 1;
                    p_pending = false;
                } else {
                    n_pap = A.apply_dot(s, pbuf[pb], Ap, part_pap_.get(), ctl);
                }
                TV* p = pbuf[pb];
                int n_rzn = -1;
                const bool rep_now = replace && ((issued + it + 1) % opt.replace_every == 0);
                if (fuse_xr && !rep_now) {
                    FusedXR<TV> fx;
                    fx.ctl = ctl;
                    fx.rz = part_rz_[cur].get();
                    fx.n_rz = n_rz;
                    fx.pap = part_pap_.get();
                    fx.n_pap = n_pap;
                    fx.x = x;
                    fx.r = r;
                    fx.p = p;
                    fx.Ap = Ap;
                    fx.rr = part_rr_.get();
                    fx.rr_cap = (int)part_rr_.n;
                    fx.n3 = n3;
                    int n_rr = 0;
                    n_rzn = M->apply_update_xr(s, fx, z, part_rz_[nxt].get(), &n_rr);
                    if (n_rzn < 0) fuse_xr = false;
                }
                if (n_rzn < 0) {
                    pcg_update_xr<TV><<<gv, kVecThreads, 0, s>>>(ctl, part_rz_[cur].get(), n_rz,
                                                                part_pap_.get(), n_pap, x, r, p, Ap, n3,
                                                                part_rr_.get());
                    SGIPC_LAUNCH_CHECK();
                    int n_rr = gv;
                    if constexpr (sizeof(TV) == 8) {
                        if (rep_now) {
                            // the true residual replaces the recursive one; the
                            // stop rule below sees its norm
                            SGIPC_HIP(hipMemcpyAsync(r, b, sizeof(TV) * n3, hipMemcpyDeviceToDevice, s));
                            opt.true_op->residual(s, b, x, r, ctl);
                            n_rr = launch_dot_partials<TV, TV>(s, r, r, n3, part_rr_.get());
                        }
                    }
                    PcgStopDesc sd{ctl, part_rz_[cur].get(), n_rz, part_rr_.get(), n_rr};
                    if (M) {
                        n_rzn = M->apply(s, r, z, part_rz_[nxt].get(), sd);
                    } else {
                        pcg_identity_apply<TV><<<gv, kVecThreads, 0, s>>>(sd, r, z, n3, part_rz_[nxt].get());
                        SGIPC_LAUNCH_CHECK();
                        n_rzn = gv;
                    }
                }
                if (n_rzn != n_rz) throw Error(SGIPC_HIP_ERROR, "preconditioner partial count changed");
                if constexpr (sizeof(TV) == 8) {
                    if (opt.record && opt.record_count && *opt.record_count < opt.record_cap &&
                        (issued + it) % (opt.record_stride > 0 ? opt.record_stride : 1) == 0) {
                        SGIPC_HIP(hipMemcpyAsync(opt.record + (size_t)(*opt.record_count) * n3, r, sizeof(TV) * n3,
                                                 hipMemcpyDeviceToDevice, s));
                        ++*opt.record_count;
                    }
                }
                if (flex) {
                    pcg_update_p_flex<TV><<<gv, kVecThreads, 0, s>>>(ctl, part_zap_.get(), n_rz, part_pap_.get(),
                                                                    n_pap, p, z, n3);
                    SGIPC_LAUNCH_CHECK();
                } else if (fuse_p) {
                    p_pending = true;
                } else {
                    pcg_update_p<TV><<<gv, kVecThreads, 0, s>>>(ctl, part_rz_[nxt].get(), part_rz_[cur].get(),
                                                               n_rz, p, z, Ap, n3);
                    SGIPC_LAUNCH_CHECK();
                }
                cur = nxt;
            }
            if (capture) {
                hipGraph_t graph = nullptr;
                guard.open = false;
                SGIPC_HIP(hipStreamEndCapture(s, &graph));
                const bool periodic = cur == cur0 && pb == pb0 && p_pending == pp0 && fuse_xr == fx0;
                if (periodic && hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) == hipSuccess) {
                    SGIPC_HIP(hipGraphDestroy(graph));
                    SGIPC_HIP(hipGraphLaunch(gexec, s));
                } else {
                    // not replayable: nothing ran during the capture, so the
                    // chunk is issued again eagerly from its start state
                    if (graph) (void)hipGraphDestroy(graph);
                    gexec = nullptr;
                    use_graph = false;
                    cur = cur0;
                    pb = pb0;
                    p_pending = pp0;
                    fuse_xr = fx0;
                    continue;
                }
            }
            issued += todo;
            ++n_chunks;
        }
        int slot = (round++) & 1;
        SGIPC_HIP(hipMemcpyAsync(host_ctl_ + slot, ctl, sizeof(PcgCtl), hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipEventRecord(ev_[slot], s));
        if (pending >= 0) {
            SGIPC_HIP(hipEventSynchronize(ev_[pending]));
            if (host_ctl_[pending].done) finished = true;
        }
        if (todo <= 0) {
            SGIPC_HIP(hipEventSynchronize(ev_[slot]));
            finished = true;
            pending = slot;
        } else {
            pending = slot;
        }
    }
    if (gexec) (void)hipGraphExecDestroy(gexec);
    SGIPC_HIP(hipStreamSynchronize(s));
    SGIPC_HIP(hipMemcpy(host_ctl_, ctl, sizeof(PcgCtl), hipMemcpyDeviceToHost));
    const PcgCtl& c = host_ctl_[0];
    res.iterations = c.iters;
    res.converged = c.done != 0;
    res.rel_residual = c.bb > 0 ? std::sqrt(c.final_rr / c.bb) : 0.0;
    res.rz = c.final_rz;
    res.rz0 = c.rz0;
    res.not_spd = c.not_spd != 0;
    res.diverged = c.diverged != 0;
    return res;
}

template class PcgSolver<double>;
template class PcgSolver<float>;

}  // namespace sgipc
