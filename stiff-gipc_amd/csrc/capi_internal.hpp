// Types shared by the C ABI translation units (not part of the public ABI).
#pragma once
#include <memory>
#include <new>
#include <string>

#include "abd.hpp"
#include "assembly.hpp"
#include "contact.hpp"
#include "mas.hpp"
#include "pcg.hpp"
#include "sgipc_internal.hpp"
#include "spmv_full.hpp"

namespace sgipc {
struct SolveWorkspace;
}

// The context owns its solver workspace and triplet converter (created on
// first use, freed by sgipc_destroy on whichever thread calls it).
struct sgipc_ctx {
    sgipc::Context c;
    std::unique_ptr<sgipc::SolveWorkspace> ws;
    std::unique_ptr<sgipc::Converter> conv;
    std::unique_ptr<sgipc::ContactDetector> contact;
    ~sgipc_ctx();
};

struct sgipc_mas {
    sgipc_ctx* ctx;
    std::unique_ptr<sgipc::Mas> m;
};

struct sgipc_spmat {
    sgipc_ctx* ctx;
    sgipc::FullMatrix m;
};

namespace sgipc {

extern thread_local std::string g_last_error;

// Run f, converting every exception into a status code (nothing aborts).
template <typename F>
int guarded(sgipc_ctx* ctx, F&& f)
{
    try {
        if (ctx) ctx->c.activate();
        f();
        return SGIPC_OK;
    } catch (const Error& e) {
        if (ctx) ctx->c.last_error = e.what();
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->c.last_error = "out of host memory";
        g_last_error = "out of host memory";
        return SGIPC_HIP_ERROR;
    } catch (const std::exception& e) {
        if (ctx) ctx->c.last_error = e.what();
        g_last_error = e.what();
        return SGIPC_HIP_ERROR;
    }
}

// Solver workspace (PCG buffers, FP32 matrix copy for refinement, the
// pinned host buffer the refinement loop reads its norms through).
struct SolveWorkspace {
    FullMatrix full;  // tiled symmetric layout of the PCG loop's SpMV
    // set while `full` already holds the layout of the matrix being solved
    // (built by the caller, e.g. concurrently with the MAS build); the PCG
    // entry then skips its own build
    bool prebuilt = false;
    SideStream layout_side;
    PcgSolver<double> pcg64;
    PcgSolver<float> pcg32;
    DevBuf<float> r32, d32;
    DevBuf<double> r64, d64;
    DevBuf<double> part;
    static constexpr int kHostSlots = 4096;
    double* host = nullptr;  // kHostSlots pinned doubles, allocated on first use
    double* pinned()
    {
        if (!host) SGIPC_HIP(hipHostMalloc(reinterpret_cast<void**>(&host), kHostSlots * sizeof(double)));
        return host;
    }
    SolveWorkspace() = default;
    SolveWorkspace(const SolveWorkspace&) = delete;
    SolveWorkspace& operator=(const SolveWorkspace&) = delete;
    ~SolveWorkspace()
    {
        if (host) (void)hipHostFree(host);
    }
};

// The subsystem preconditioner of GlobalLinearSystem::apply_preconditioner
// (global_linear_system.cu:150-220): z = r, block-Jacobi on the affine-body
// rows [0, 4 nb), MAS on the FEM rows [mas_row_offset, n).
struct SubsystemPrec {
    Mas* mas = nullptr;
    int mas_row_offset = 0;
    const AbdJacobi* abd = nullptr;
};

template <typename TV>
struct SubsystemPrecond : Preconditioner<TV> {
    SubsystemPrec p;
    explicit SubsystemPrecond(const SubsystemPrec& q) : p(q) {}
    int apply(hipStream_t s, const TV* r, TV* z, double* rz_part, const PcgStopDesc& sd) override
    {
        int k = 0;
        // the MAS evaluates the stop rule in its first kernel; the body
        // blocks run after it and skip once the flag is set
        if (p.mas)
            k = p.mas->template apply<TV>(s, r + 3L * p.mas_row_offset, z + 3L * p.mas_row_offset, rz_part, sd);
        if (p.abd && p.abd->n_bodies > 0)
            k += p.abd->template apply<TV>(s, r, z, rz_part ? rz_part + k : nullptr, sd.ctl);
        return k;
    }
    // the fused x/r update covers the MAS rows only: offered when they are
    // the whole system
    int apply_update_xr(hipStream_t s, const FusedXR<TV>& fx, TV* z, double* rz_part, int* n_rr) override
    {
        if (!p.mas || p.mas_row_offset != 0 || (p.abd && p.abd->n_bodies > 0)) return -1;
        return p.mas->template apply_update_xr<TV>(s, fx, z, rz_part, n_rr);
    }
    int max_partials() const override { return 1024 + (p.abd ? p.abd->partials() : 0); }
    int max_rr_partials() const override { return p.mas ? p.mas->rr_partials() : 0; }
    // flexible PCG needs z.Ap over every row: only when the MAS covers them
    bool set_flex(const TV* Ap, double* zap) override
    {
        if (!p.mas || p.mas_row_offset != 0 || (p.abd && p.abd->n_bodies > 0)) return false;
        p.mas->flex_Ap_ = Ap;
        p.mas->flex_zap_ = zap;
        return true;
    }
};

// resolves the allocation-free scratch array's address on `device` (current)
void noalloc_setup(int device);

// learned-level helpers (capi_mpnn.cpp)
sgipc_mpnn_data* mpnn_data_new(sgipc_ctx* ctx);
CoarseGraph& mpnn_data_graph(sgipc_mpnn_data* d);
void mpnn_data_append(sgipc_mpnn_data* d, const float* X, const float* q, int n_samples);
Mpnn* mpnn_of(sgipc_mpnn* h);

int pcg_solve_impl(Context& c, SolveWorkspace& ws, const double* vals, const int* rows, const int* cols,
                   int nnz, int n_rows, const SubsystemPrec& prec, const double* b, double* x,
                   const sgipc_pcg_opts& o, sgipc_pcg_stats* stats, const PcgOptions* extra = nullptr);

}  // namespace sgipc
