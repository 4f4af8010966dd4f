// C ABI over the gfx950 kernels (include/sgipc.h): context, device memory
// helpers, SpMV, MAS and PCG (FP64, reference FP32-MAS, and FP32 PCG inside
// FP64 iterative refinement).  Every entry point catches sgipc::Error /
// std::exception and turns it into a status code; nothing aborts the process.
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "capi_internal.hpp"

namespace sgipc {

thread_local std::string g_last_error;

// PCG operator over the expanded (both triangles) matrix: Ap = A p written
// row by row (no zero-fill needed) with the p.Ap partials.
template <typename TV>
struct FullOp : SpmvOp<TV> {
    const FullMatrix* m;
    bool f32_values = false;  // FP64 vectors against the FP32 matrix copy
    int apply_dot(hipStream_t s, const TV* p, TV* Ap, double* part, const PcgCtl* ctl) const override
    {
        if constexpr (sizeof(TV) == 8)
            if (f32_values) return m->apply_f32_values(s, 1.0, p, 0.0, Ap, part, ctl ? &ctl->done : nullptr);
        return m->apply<TV>(s, TV(1), p, TV(0), Ap, part, ctl ? &ctl->done : nullptr);
    }
    bool fuses_update_p() const override { return true; }
    // r holds b on entry: r = b - A x with the FP64 values
    bool residual(hipStream_t s, const TV*, const TV* x, TV* r, const PcgCtl* ctl) const override
    {
        if constexpr (sizeof(TV) == 8) {
            m->apply<double>(s, -1.0, x, 1.0, r, nullptr, ctl ? &ctl->done : nullptr);
            return true;
        }
        return false;
    }
    int apply_dot_update_p(hipStream_t s, const TV* z, const TV* p_old, TV* p_new, TV* Ap, double* part,
                           const PcgCtl* ctl, const double* rz_new, const double* rz_old, int n_rz) const override
    {
        return m->apply_update_p<TV>(s, f32_values, z, p_old, p_new, Ap, part, ctl ? &ctl->done : nullptr, rz_new,
                                     rz_old, n_rz);
    }
};

// r = b - A x in FP64 with ||r||^2 partials; returns the partial count.
int residual(hipStream_t s, const FullMatrix& A, const double* x, const double* b, double* r, double* part, long n3)
{
    SGIPC_HIP(hipMemcpyAsync(r, b, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
    A.apply<double>(s, -1.0, x, 1.0, r, nullptr, nullptr);
    return launch_dot_partials<double, double>(s, r, r, n3, part);
}

}  // namespace sgipc

sgipc_ctx::~sgipc_ctx() = default;

using namespace sgipc;

extern "C" {

int sgipc_abi_version(void) { return SGIPC_ABI_VERSION; }

int sgipc_create(sgipc_ctx** out, int device)
{
    if (!out) return SGIPC_BAD_ARG;
    *out = nullptr;
    auto* ctx = new (std::nothrow) sgipc_ctx();
    if (!ctx) return SGIPC_HIP_ERROR;
    int st = guarded(nullptr, [&] {
        int count = 0;
        SGIPC_HIP(hipGetDeviceCount(&count));
        SGIPC_CHECK(device >= 0 && device < count, SGIPC_BAD_ARG, "no such HIP device");
        ctx->c.device = device;
        SGIPC_HIP(hipSetDevice(device));
        SGIPC_HIP(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
        ctx->c.own_stream = true;
        hipDeviceProp_t prop;
        SGIPC_HIP(hipGetDeviceProperties(&prop, device));
        ctx->c.num_cus = prop.multiProcessorCount;
        noalloc_setup(device);
    });
    if (st != SGIPC_OK) {
        delete ctx;
        return st;
    }
    *out = ctx;
    return SGIPC_OK;
}

int sgipc_destroy(sgipc_ctx* ctx)
{
    if (!ctx) return SGIPC_BAD_ARG;
    int st = guarded(ctx, [&] {
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
        ctx->ws.reset();
        ctx->conv.reset();
        ctx->contact.reset();
        if (ctx->c.own_stream && ctx->c.stream) SGIPC_HIP(hipStreamDestroy(ctx->c.stream));
    });
    delete ctx;
    return st;
}

int sgipc_set_stream(sgipc_ctx* ctx, void* stream)
{
    if (!ctx) return SGIPC_BAD_ARG;
    return guarded(ctx, [&] {
        if (ctx->c.own_stream && ctx->c.stream) SGIPC_HIP(hipStreamDestroy(ctx->c.stream));
        ctx->c.stream = static_cast<hipStream_t>(stream);
        ctx->c.own_stream = false;
    });
}

int sgipc_synchronize(sgipc_ctx* ctx)
{
    if (!ctx) return SGIPC_BAD_ARG;
    return guarded(ctx, [&] { SGIPC_HIP(hipStreamSynchronize(ctx->c.stream)); });
}

const char* sgipc_last_error(const sgipc_ctx* ctx)
{
    if (ctx) return ctx->c.last_error.c_str();
    return g_last_error.c_str();
}

int sgipc_malloc(sgipc_ctx* ctx, void** ptr, size_t bytes)
{
    if (!ctx || !ptr) return SGIPC_BAD_ARG;
    return guarded(ctx, [&] { SGIPC_HIP(hipMalloc(ptr, bytes ? bytes : 1)); });
}

int sgipc_free(sgipc_ctx* ctx, void* ptr)
{
    if (!ctx) return SGIPC_BAD_ARG;
    return guarded(ctx, [&] {
        if (ptr) SGIPC_HIP(hipFree(ptr));
    });
}

int sgipc_memcpy_htod(sgipc_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !src))) return SGIPC_BAD_ARG;
    return guarded(ctx, [&] {
        if (bytes) SGIPC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->c.stream));
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int sgipc_memcpy_dtoh(sgipc_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !src))) return SGIPC_BAD_ARG;
    return guarded(ctx, [&] {
        if (bytes) SGIPC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->c.stream));
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int sgipc_memcpy_dtod(sgipc_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !src))) return SGIPC_BAD_ARG;
    return guarded(ctx, [&] {
        if (bytes) SGIPC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->c.stream));
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int sgipc_memset(sgipc_ctx* ctx, void* dst, int value, size_t bytes)
{
    if (!ctx || (bytes && !dst)) return SGIPC_BAD_ARG;
    return guarded(ctx, [&] {
        if (bytes) SGIPC_HIP(hipMemsetAsync(dst, value, bytes, ctx->c.stream));
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

// ---------------------------------------------------------------- SpMV
int sgipc_spmv_srbk(sgipc_ctx* ctx, double a, const double* vals, const int* rows, const int* cols,
                    int nnz, const double* x, double b, double* y, int n_rows)
{
    if (!ctx) return SGIPC_BAD_ARG;
    if (n_rows < 0 || nnz < 0) return SGIPC_BAD_ARG;
    if (n_rows == 0) return SGIPC_EMPTY;
    if ((nnz > 0 && (!vals || !rows || !cols || !x)) || !y) return SGIPC_BAD_ARG;
    return guarded(ctx, [&] {
        launch_spmv_f64(ctx->c.stream, a, vals, rows, cols, nnz, x, b, y, n_rows, nullptr, nullptr);
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int sgipc_spmv_srbk_f32(sgipc_ctx* ctx, float a, const float* vals, const int* rows, const int* cols,
                        int nnz, const float* x, float b, float* y, int n_rows)
{
    if (!ctx) return SGIPC_BAD_ARG;
    if (n_rows < 0 || nnz < 0) return SGIPC_BAD_ARG;
    if (n_rows == 0) return SGIPC_EMPTY;
    if ((nnz > 0 && (!vals || !rows || !cols || !x)) || !y) return SGIPC_BAD_ARG;
    return guarded(ctx, [&] {
        launch_spmv_f32(ctx->c.stream, a, vals, rows, cols, nnz, x, b, y, n_rows);
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int sgipc_spmat_create(sgipc_ctx* ctx, sgipc_spmat** out, const double* vals, const int* rows, const int* cols,
                       int nnz, int n_rows, int want_f32)
{
    if (!ctx || !out) return SGIPC_BAD_ARG;
    *out = nullptr;
    if (nnz < 0 || n_rows < 0) return SGIPC_BAD_ARG;
    if (nnz == 0 || n_rows == 0) return SGIPC_EMPTY;
    if (!vals || !rows || !cols) return SGIPC_BAD_ARG;
    auto* m = new (std::nothrow) sgipc_spmat();
    if (!m) return SGIPC_HIP_ERROR;
    m->ctx = ctx;
    int st = guarded(ctx, [&] {
        m->m.build(ctx->c.stream, vals, rows, cols, nnz, n_rows, want_f32 != 0);
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
    if (st != SGIPC_OK) {
        delete m;
        return st;
    }
    *out = m;
    return SGIPC_OK;
}

int sgipc_spmat_destroy(sgipc_spmat* m)
{
    if (!m) return SGIPC_BAD_ARG;
    int st = guarded(m->ctx, [&] { SGIPC_HIP(hipStreamSynchronize(m->ctx->c.stream)); });
    delete m;
    return st;
}

int sgipc_spmat_apply(sgipc_spmat* m, double a, const double* x, double b, double* y)
{
    if (!m || !x || !y) return SGIPC_BAD_ARG;
    return guarded(m->ctx, [&] {
        m->m.apply<double>(m->ctx->c.stream, a, x, b, y, nullptr, nullptr);
        SGIPC_HIP(hipStreamSynchronize(m->ctx->c.stream));
    });
}

int sgipc_spmat_apply_f32(sgipc_spmat* m, float a, const float* x, float b, float* y)
{
    if (!m || !x || !y) return SGIPC_BAD_ARG;
    return guarded(m->ctx, [&] {
        m->m.apply<float>(m->ctx->c.stream, a, x, b, y, nullptr, nullptr);
        SGIPC_HIP(hipStreamSynchronize(m->ctx->c.stream));
    });
}

int sgipc_spmat_apply_f32_values(sgipc_spmat* m, double a, const double* x, double b, double* y)
{
    if (!m || !x || !y) return SGIPC_BAD_ARG;
    return guarded(m->ctx, [&] {
        m->m.apply_f32_values(m->ctx->c.stream, a, x, b, y, nullptr, nullptr);
        SGIPC_HIP(hipStreamSynchronize(m->ctx->c.stream));
    });
}

int sgipc_spmat_layout(const sgipc_spmat* m, int* halo_form, int* halo_max, int* tl_slots)
{
    if (!m) return SGIPC_BAD_ARG;
    if (halo_form) *halo_form = m->m.use_halo ? 1 : 0;
    if (halo_max) *halo_max = m->m.halo_max;
    if (tl_slots) *tl_slots = m->m.tl_max;
    return SGIPC_OK;
}

int sgipc_spmat_long_rows(const sgipc_spmat* m, int* n_long)
{
    if (!m || !n_long) return SGIPC_BAD_ARG;
    *n_long = m->m.n_long;
    return SGIPC_OK;
}

int sgipc_spmv_time(sgipc_ctx* ctx, int precision, const double* vals, const int* rows, const int* cols, int nnz,
                    int n_rows, int reps, int cold_cache, double* avg_ms, int* stored_blocks)
{
    if (!ctx || !vals || !rows || !cols || !avg_ms || reps < 1 || nnz < 1 || n_rows < 1) return SGIPC_BAD_ARG;
    if (precision != SGIPC_FP64 && precision != SGIPC_FP32_IR) return SGIPC_BAD_ARG;
    return guarded(ctx, [&] {
        hipStream_t s = ctx->c.stream;
        FullMatrix A;
        A.build(s, vals, rows, cols, nnz, n_rows, precision == SGIPC_FP32_IR);
        if (stored_blocks) *stored_blocks = (int)A.stored_blocks(s);
        const long n3 = 3L * n_rows;
        // FP32_IR times the PCG loop's kernel: FP32 matrix values, FP64 vectors
        DevBuf<double> x64, y64, part, flush;
        part.reserve(A.partials() > 0 ? A.partials() : 1);
        x64.reserve(n3);
        y64.reserve(n3);
        launch_fill<double>(s, x64.get(), 1.0, n3);
        // 512 MiB, twice the Infinity Cache: streaming it evicts the matrix
        const long n_flush = cold_cache ? (512L << 20) / (long)sizeof(double) : 0;
        if (cold_cache) {
            flush.reserve(n_flush);
            launch_fill<double>(s, flush.get(), 0.0, n_flush);
        }
        auto launch = [&] {
            if (precision == SGIPC_FP64)
                A.apply<double>(s, 1.0, x64.get(), 0.0, y64.get(), part.get(), nullptr);
            else
                A.apply_f32_values(s, 1.0, x64.get(), 0.0, y64.get(), part.get(), nullptr);
        };
        launch();  // warm-up
        std::vector<hipEvent_t> ev(2 * (size_t)reps);
        for (auto& e : ev) SGIPC_HIP(hipEventCreate(&e));
        // the events ride on the SpMV's own dispatch: they bracket the kernel
        // itself, as a kernel trace does
        for (int i = 0; i < reps; ++i) {
            if (cold_cache) launch_stream_read(s, flush.get(), n_flush, part.get());
            A.time_start = ev[2 * i];
            A.time_stop = ev[2 * i + 1];
            launch();
        }
        A.time_start = A.time_stop = nullptr;
        SGIPC_HIP(hipEventSynchronize(ev.back()));
        double total = 0.0;
        for (int i = 0; i < reps; ++i) {
            float ms = 0.f;
            SGIPC_HIP(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
            total += ms;
        }
        *avg_ms = total / reps;
        for (auto& e : ev) (void)hipEventDestroy(e);
    });
}

int sgipc_convert_f64_to_f32(sgipc_ctx* ctx, const double* in, float* out, size_t n)
{
    if (!ctx || (n && (!in || !out))) return SGIPC_BAD_ARG;
    return guarded(ctx, [&] {
        launch_f64_to_f32(ctx->c.stream, in, out, n);
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

// ---------------------------------------------------------------- MAS
int sgipc_mas_create(sgipc_ctx* ctx, sgipc_mas** out, int bank_size)
{
    if (!ctx || !out) return SGIPC_BAD_ARG;
    *out = nullptr;
    if (bank_size != 16 && bank_size != 32) return SGIPC_BAD_ARG;
    auto* m = new (std::nothrow) sgipc_mas();
    if (!m) return SGIPC_HIP_ERROR;
    m->ctx = ctx;
    int st = guarded(ctx, [&] { m->m = std::make_unique<Mas>(&ctx->c, bank_size); });
    if (st != SGIPC_OK) {
        delete m;
        return st;
    }
    *out = m;
    return SGIPC_OK;
}

int sgipc_mas_destroy(sgipc_mas* mas)
{
    if (!mas) return SGIPC_BAD_ARG;
    int st = guarded(mas->ctx, [&] { mas->m.reset(); });
    delete mas;
    return st;
}

int sgipc_mas_init_neighbor(sgipc_mas* mas, int vert_num, int collision_node_offset,
                            int total_neighbor_num, const uint32_t* nbr_start, const uint32_t* nbr_num,
                            const uint32_t* nbr_list)
{
    if (!mas) return SGIPC_BAD_ARG;
    if (vert_num < 1) return SGIPC_EMPTY;
    if (total_neighbor_num < 0) return SGIPC_BAD_ARG;
    return guarded(mas->ctx, [&] {
        mas->m->init_neighbor(vert_num, collision_node_offset, total_neighbor_num, nbr_start, nbr_num, nbr_list);
    });
}

int sgipc_mas_init_matrix(sgipc_mas* mas)
{
    if (!mas) return SGIPC_BAD_ARG;
    return guarded(mas->ctx, [&] { mas->m->init_matrix(); });
}

int sgipc_mas_set_preconditioner_bcoo(sgipc_mas* mas, const double* vals, const int* rows, const int* cols,
                                      int nnz, int offset, const int* pairs, int n_pairs, int precision)
{
    if (!mas) return SGIPC_BAD_ARG;
    if (nnz < 0 || n_pairs < 0) return SGIPC_BAD_ARG;
    if (nnz > 0 && (!vals || !rows || !cols)) return SGIPC_BAD_ARG;
    return guarded(mas->ctx, [&] {
        mas->m->set_bcoo(vals, rows, cols, nnz, offset, pairs, n_pairs, precision);
    });
}

int sgipc_mas_preconditioning(sgipc_mas* mas, const double* r, double* z)
{
    if (!mas || !r || !z) return SGIPC_BAD_ARG;
    return guarded(mas->ctx, [&] {
        hipStream_t s = mas->ctx->c.stream;
        PcgStopDesc none{};
        mas->m->apply<double>(s, r, z, nullptr, none);
        SGIPC_HIP(hipStreamSynchronize(s));
    });
}

int sgipc_mas_set_coarse_mode(sgipc_mas* mas, int mode)
{
    if (!mas || mode < 0 || mode > 2) return SGIPC_BAD_ARG;
    mas->m->coarse_mode = mode;
    return SGIPC_OK;
}

int sgipc_mas_set_smoothing_weight(sgipc_mas* mas, double omega)
{
    if (!mas || !(omega > 0.0) || !(omega < 2.0)) return SGIPC_BAD_ARG;
    mas->m->omega0 = omega;
    return SGIPC_OK;
}

int sgipc_mas_set_level0_half(sgipc_mas* mas, int on)
{
    if (!mas || on < 0 || on > 1) return SGIPC_BAD_ARG;
    mas->m->half_level0 = on != 0;
    return SGIPC_OK;
}

int sgipc_mas_time_apply(sgipc_mas* mas, const double* r, double* z, int reps, double* avg_ms)
{
    if (!mas || !r || !z || !avg_ms || reps < 1) return SGIPC_BAD_ARG;
    return guarded(mas->ctx, [&] {
        hipStream_t s = mas->ctx->c.stream;
        PcgStopDesc none{};
        hipEvent_t e0, e1;
        SGIPC_HIP(hipEventCreate(&e0));
        SGIPC_HIP(hipEventCreate(&e1));
        mas->m->apply<double>(s, r, z, nullptr, none);
        SGIPC_HIP(hipEventRecord(e0, s));
        for (int i = 0; i < reps; ++i) mas->m->apply<double>(s, r, z, nullptr, none);
        SGIPC_HIP(hipEventRecord(e1, s));
        SGIPC_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        SGIPC_HIP(hipEventElapsedTime(&ms, e0, e1));
        *avg_ms = ms / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

int sgipc_mas_levels(sgipc_mas* mas, int* level_num, int* offsets, int* counts)
{
    if (!mas || !level_num || !offsets || !counts) return SGIPC_BAD_ARG;
    return guarded(mas->ctx, [&] { *level_num = mas->m->levels(offsets, counts); });
}

int sgipc_mas_tables(sgipc_mas* mas, int* coarse_table, int* going_next)
{
    if (!mas) return SGIPC_BAD_ARG;
    return guarded(mas->ctx, [&] { mas->m->tables(coarse_table, going_next); });
}

int sgipc_mas_inverse_blocks(sgipc_mas* mas, double* out)
{
    if (!mas || !out) return SGIPC_BAD_ARG;
    return guarded(mas->ctx, [&] {
        SGIPC_CHECK(mas->m->built, SGIPC_BAD_ARG, "MAS not set");
        mas->m->inverse_blocks(out);
    });
}

int sgipc_mas_level2_inverse(sgipc_mas* mas, double* out, int* m)
{
    if (!mas || !m) return SGIPC_BAD_ARG;
    *m = 0;
    return guarded(mas->ctx, [&] {
        SGIPC_CHECK(mas->m->built, SGIPC_BAD_ARG, "MAS not set");
        *m = mas->m->level2_inverse(out);
    });
}

// ---------------------------------------------------------------- PCG
void sgipc_pcg_default_opts(sgipc_pcg_opts* o)
{
    if (!o) return;
    o->tol = 1e-5;
    o->stop_mode = SGIPC_STOP_TRUE_RESIDUAL;
    o->max_iter = 0;
    o->precision = SGIPC_FP64;
    o->check_every = 8;
    o->fuse = 1;
    o->flexible = 0;
    o->ir_replace_every = 8;
    o->graph = 0;
    o->ir_max_outer = 20;
    o->ir_inner_tol = 0.0;
}

}  // extern "C"

namespace sgipc {

// Shared by sgipc_pcg_solve and the linear-system object.
int pcg_solve_impl(Context& c, SolveWorkspace& ws, const double* vals, const int* rows, const int* cols,
                   int nnz, int n_rows, const SubsystemPrec& prec, const double* b, double* x,
                   const sgipc_pcg_opts& o, sgipc_pcg_stats* stats, const PcgOptions* extra)
{
    const bool has_prec = prec.mas != nullptr || (prec.abd != nullptr && prec.abd->n_bodies > 0);
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = c.stream;
    const long n3 = 3L * n_rows;
    int max_iter = o.max_iter > 0 ? o.max_iter : (int)(0.3 * (double)n3);
    if (max_iter < 2) max_iter = 2;
    sgipc_pcg_stats st{};
    int status = SGIPC_OK;
    ws.part.reserve(4096);
    // coarse mode 2 runs the PCG matrix inside the MAS apply: only when the
    // MAS covers every row; an FP32 MAS apply reads its FP32 copy
    const bool mas_fine = prec.mas != nullptr && prec.mas->coarse_mode == 2 && prec.mas_row_offset == 0 &&
                          !(prec.abd != nullptr && prec.abd->n_bodies > 0) && prec.mas->n == n_rows;
    const bool want_f32 = o.precision == SGIPC_FP32_IR || (mas_fine && prec.mas->precision != SGIPC_FP64);
    if (!(ws.prebuilt && ws.full.n_rows == n_rows && ws.full.nnz == nnz && (ws.full.has_f32 || !want_f32)))
        ws.full.build(s, vals, rows, cols, nnz, n_rows, want_f32);
    if (prec.mas != nullptr && (mas_fine || !prec.mas->build_own_fine)) prec.mas->fine_ = mas_fine ? &ws.full : nullptr;
    if (o.precision != SGIPC_FP32_IR) {
        FullOp<double> A;
        A.nnz = ws.full.nnz;
        A.max_partials = ws.full.partials();
        A.m = &ws.full;
        SubsystemPrecond<double> M(prec);
        PcgOptions po;
        po.tol = o.tol;
        po.stop_mode = o.stop_mode;
        po.max_iter = max_iter;
        po.check_every = o.check_every;
        po.fuse = o.fuse;
        po.graph = o.graph != 0;
        po.flexible = o.flexible != 0;
        if (extra) {
            po.record = extra->record;
            po.record_stride = extra->record_stride;
            po.record_cap = extra->record_cap;
            po.record_count = extra->record_count;
            po.flexible = po.flexible || extra->flexible;
        }
        PcgResult r = ws.pcg64.solve(s, A, has_prec ? &M : nullptr, b, x, n3, po);
        st.iterations = r.iterations;
        st.outer_iterations = 1;
        st.rz = r.rz;
        st.rz0 = r.rz0;
        if (r.not_spd) status = SGIPC_NOT_SPD;
        else if (!r.converged) status = SGIPC_MAXITER;
    } else {
        // FP64 iterative refinement: the inner PCG keeps FP64 vectors, dot
        // products and alpha/beta, and reads the FP32 copies of the matrix
        // (SpMV) and of the MAS inverse blocks (apply); the outer residual
        // b - A x uses the FP64 matrix.  Inside a pass the recursive residual
        // is replaced by the FP64 one every ir_replace_every iterations, so
        // one Krylov run usually meets the tolerance; a further pass (a
        // restart) only follows when it did not.
        ws.r64.reserve(n3);
        ws.d64.reserve(n3);
        FullOp<double> A;
        A.nnz = ws.full.nnz;
        A.max_partials = ws.full.partials();
        A.m = &ws.full;
        A.f32_values = true;
        FullOp<double> A64;
        A64.nnz = ws.full.nnz;
        A64.max_partials = ws.full.partials();
        A64.m = &ws.full;
        SubsystemPrecond<double> M(prec);
        launch_fill<double>(s, x, 0.0, n3);
        SGIPC_HIP(hipMemcpyAsync(ws.r64.get(), b, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
        double* hp = ws.pinned();
        auto host_sum = [&](const double* dpart, int np) {
            SGIPC_HIP(hipMemcpyAsync(hp, dpart, sizeof(double) * np, hipMemcpyDeviceToHost, s));
            SGIPC_HIP(hipStreamSynchronize(s));
            double acc = 0.0;
            for (int i = 0; i < np; ++i) acc += hp[i];
            return acc;
        };
        int np = launch_dot_partials<double, double>(s, b, b, n3, ws.part.get());
        const double bb = host_sum(ws.part.get(), np);
        double rr = bb;
        int total_it = 0, outer = 0;
        bool ok = bb == 0.0;
        const int max_outer = o.ir_max_outer > 0 ? o.ir_max_outer : 20;
        double best_rr = rr;
        while (!ok && outer < max_outer && total_it < max_iter) {
            ++outer;
            PcgOptions po;
            // reduce ||r|| by the factor still missing (at least by half)
            double need = o.tol * std::sqrt(bb / rr) * 0.8;
            if (o.ir_inner_tol > 0 && need < o.ir_inner_tol) need = o.ir_inner_tol;
            if (need > 0.5) need = 0.5;
            po.tol = need;
            po.stop_mode = SGIPC_STOP_TRUE_RESIDUAL;
            po.max_iter = max_iter - total_it;
            po.check_every = o.check_every;
            po.fuse = o.fuse;
            po.graph = o.graph != 0;
            po.flexible = o.flexible != 0 || (extra && extra->flexible);
            po.true_op = &A64;
            po.replace_every = o.ir_replace_every;
            // a pass whose FP32 operators cannot resolve the system ends once
            // its FP64 residual has grown 1e4 times past the pass's start
            po.diverge = 1e8;
            PcgResult r = ws.pcg64.solve(s, A, has_prec ? &M : nullptr, ws.r64.get(), ws.d64.get(), n3, po);
            total_it += r.iterations;
            if (r.not_spd) {
                status = SGIPC_NOT_SPD;
                break;
            }
            launch_axpy_f64(s, 1.0, ws.d64.get(), x, n3);
            int nrr = residual(s, ws.full, x, b, ws.r64.get(), ws.part.get(), n3);
            rr = host_sum(ws.part.get(), nrr);
            if (rr <= o.tol * o.tol * bb) ok = true;
            if (!ok && rr >= best_rr) {
                // the pass made the FP64 residual worse (the FP32 operators
                // cannot resolve this system): take its correction back and
                // stop with the best iterate
                launch_axpy_f64(s, -1.0, ws.d64.get(), x, n3);
                nrr = residual(s, ws.full, x, b, ws.r64.get(), ws.part.get(), n3);
                rr = host_sum(ws.part.get(), nrr);
                break;
            }
            if (rr > best_rr * 0.999999 && outer > 1) break;  // stagnation guard
            if (rr < best_rr) best_rr = rr;
        }
        st.iterations = total_it;
        st.outer_iterations = outer;
        if (status == SGIPC_OK && !ok) status = SGIPC_MAXITER;
    }
    // true FP64 residual of the returned x
    {
        ws.r64.reserve(n3);
        int nrr = residual(s, ws.full, x, b, ws.r64.get(), ws.part.get(), n3);
        std::vector<double> h(nrr);
        SGIPC_HIP(hipMemcpyAsync(h.data(), ws.part.get(), sizeof(double) * nrr, hipMemcpyDeviceToHost, s));
        int nbb = launch_dot_partials<double, double>(s, b, b, n3, ws.part.get() + 2048);
        std::vector<double> hb(nbb);
        SGIPC_HIP(hipMemcpyAsync(hb.data(), ws.part.get() + 2048, sizeof(double) * nbb, hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipStreamSynchronize(s));
        double rr = 0, bb = 0;
        for (double v : h) rr += v;
        for (double v : hb) bb += v;
        st.rel_residual = bb > 0 ? std::sqrt(rr / bb) : 0.0;
        if (o.stop_mode == SGIPC_STOP_TRUE_RESIDUAL && status == SGIPC_MAXITER && st.rel_residual <= o.tol)
            status = SGIPC_OK;
    }
    st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    st.status = status;
    if (stats) *stats = st;
    return status;
}

}  // namespace sgipc

extern "C" int sgipc_pcg_solve(sgipc_ctx* ctx, const double* vals, const int* rows, const int* cols, int nnz,
                               int n_rows, sgipc_mas* mas, const double* b, double* x,
                               const sgipc_pcg_opts* opts, sgipc_pcg_stats* stats)
{
    if (!ctx || !b || !x) return SGIPC_BAD_ARG;
    if (n_rows <= 0 || nnz <= 0) return SGIPC_EMPTY;
    if (!vals || !rows || !cols) return SGIPC_BAD_ARG;
    sgipc_pcg_opts o;
    sgipc_pcg_default_opts(&o);
    if (opts) o = *opts;
    if (o.precision < SGIPC_FP64 || o.precision > SGIPC_FP32_IR) return SGIPC_BAD_ARG;
    if (o.stop_mode != 0 && o.stop_mode != 1) return SGIPC_BAD_ARG;
    if (mas && mas->m->n != n_rows) return SGIPC_BAD_ARG;
    int status = SGIPC_OK;
    int st = guarded(ctx, [&] {
        auto& ws = ctx->ws;
        if (!ws) ws = std::make_unique<SolveWorkspace>();
        SubsystemPrec prec;
        prec.mas = mas ? mas->m.get() : nullptr;
        status = pcg_solve_impl(ctx->c, *ws, vals, rows, cols, nnz, n_rows, prec, b, x, o,
                                stats);
    });
    return st != SGIPC_OK ? st : status;
}
