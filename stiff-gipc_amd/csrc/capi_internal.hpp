// Types shared by the C ABI translation units (not part of the public ABI).
#pragma once
#include <memory>
#include <new>
#include <string>

#include "assembly.hpp"
#include "mas.hpp"
#include "pcg.hpp"
#include "sgipc_internal.hpp"
#include "spmv_full.hpp"

struct sgipc_ctx {
    sgipc::Context c;
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

// Solver workspace (PCG buffers, FP32 matrix copy for refinement).
struct SolveWorkspace {
    FullMatrix full;  // tiled symmetric layout of the PCG loop's SpMV
    PcgSolver<double> pcg64;
    PcgSolver<float> pcg32;
    DevBuf<float> r32, d32;
    DevBuf<double> r64, d64;
    DevBuf<double> part;
};

int pcg_solve_impl(Context& c, SolveWorkspace& ws, const double* vals, const int* rows, const int* cols,
                   int nnz, int n_rows, Mas* mas, const double* b, double* x, const sgipc_pcg_opts& o,
                   sgipc_pcg_stats* stats);

}  // namespace sgipc
