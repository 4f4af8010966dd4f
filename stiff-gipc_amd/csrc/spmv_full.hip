// SELL-64-sigma symmetric SpMV for the PCG loop on gfx950.
//
// The reference streams the upper BCOO and scatters the transposed half with
// atomics (Spmv::srbk_spmv, StiffGIPC/linear_system/utils/spmv.cu:188-313).
// For the hundreds of SpMVs of one solve the matrix is instead expanded once
// into both triangles (the transposed half ordered by a stable (col,row)
// sort, so the layout is deterministic) and packed as SELL-64-sigma:
//   * rows are stably sorted by length (longest first), 64 rows per slice;
//     the elastic rows all have 15 blocks, so slices are almost unpadded and
//     the surface rows that carry contact blocks gather in a few wide slices;
//   * a slice's blocks are stored slot by slot, component-major: value c of
//     slot s of lane l sits at ((S_q + s) * 9 + c) * 64 + l, so each of the 9
//     loads of a slot is one contiguous 512-byte wavefront access;
//   * one lane owns one row and accumulates it in registers: no LDS, no
//     shuffles, no atomics, no zero-fill of y; every row is written once.
// x^T(Ax) partials for PCG are formed on the way.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "spmv_full.hpp"
#include "wave_ops.hpp"

namespace sgipc {

namespace {

constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kMaxGrid = 2048;

template <typename TM, typename TV, bool kDot>
__global__ __launch_bounds__(kThreads) void k_spmv_sell(const TM* __restrict__ vals, const int* __restrict__ cols,
                                                        const int* __restrict__ slice_ptr,
                                                        const int* __restrict__ slice_w, const int* __restrict__ perm,
                                                        int n_rows, int n_slices, TV a, const TV* __restrict__ x,
                                                        TV b, TV* __restrict__ y, double* __restrict__ dot_part,
                                                        const int* __restrict__ done)
{
    __shared__ double red[16];
    if (done != nullptr && *done) return;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    double dot = 0.0;
    for (int q = blockIdx.x * kWaves + w; q < n_slices; q += gridDim.x * kWaves) {
        const int p = q * 64 + lane;
        const int i = p < n_rows ? perm[p] : -1;
        const long S = slice_ptr[q];
        const int W = slice_w[q];
        TV acc0 = 0, acc1 = 0, acc2 = 0;
        const int* cp = cols + S * 64 + lane;
        const TM* vp = vals + S * 9 * 64 + lane;
        int s = 0;
        for (; s + 2 <= W; s += 2) {
            const int j0 = cp[(long)s * 64];
            const int j1 = cp[(long)(s + 1) * 64];
            TM h0[9], h1[9];
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                h0[c] = vp[((long)s * 9 + c) * 64];
                h1[c] = vp[((long)(s + 1) * 9 + c) * 64];
            }
            const TV a0 = x[3 * (long)j0], a1 = x[3 * (long)j0 + 1], a2 = x[3 * (long)j0 + 2];
            const TV b0 = x[3 * (long)j1], b1 = x[3 * (long)j1 + 1], b2 = x[3 * (long)j1 + 2];
            acc0 += (TV)h0[0] * a0 + (TV)h0[3] * a1 + (TV)h0[6] * a2;
            acc1 += (TV)h0[1] * a0 + (TV)h0[4] * a1 + (TV)h0[7] * a2;
            acc2 += (TV)h0[2] * a0 + (TV)h0[5] * a1 + (TV)h0[8] * a2;
            acc0 += (TV)h1[0] * b0 + (TV)h1[3] * b1 + (TV)h1[6] * b2;
            acc1 += (TV)h1[1] * b0 + (TV)h1[4] * b1 + (TV)h1[7] * b2;
            acc2 += (TV)h1[2] * b0 + (TV)h1[5] * b1 + (TV)h1[8] * b2;
        }
        if (s < W) {
            const int j0 = cp[(long)s * 64];
            TM h0[9];
#pragma unroll
            for (int c = 0; c < 9; ++c) h0[c] = vp[((long)s * 9 + c) * 64];
            const TV a0 = x[3 * (long)j0], a1 = x[3 * (long)j0 + 1], a2 = x[3 * (long)j0 + 2];
            acc0 += (TV)h0[0] * a0 + (TV)h0[3] * a1 + (TV)h0[6] * a2;
            acc1 += (TV)h0[1] * a0 + (TV)h0[4] * a1 + (TV)h0[7] * a2;
            acc2 += (TV)h0[2] * a0 + (TV)h0[5] * a1 + (TV)h0[8] * a2;
        }
        if (i >= 0) {
            TV* yi = y + 3 * (long)i;
            if (b == TV(0)) {
                yi[0] = a * acc0;
                yi[1] = a * acc1;
                yi[2] = a * acc2;
            } else {
                yi[0] = a * acc0 + b * yi[0];
                yi[1] = a * acc1 + b * yi[1];
                yi[2] = a * acc2 + b * yi[2];
            }
            if (kDot)
                dot += (double)x[3 * (long)i] * (double)acc0 + (double)x[3 * (long)i + 1] * (double)acc1 +
                       (double)x[3 * (long)i + 2] * (double)acc2;
        }
    }
    if (kDot) {
        double sblk = block_sum(dot, red);
        if (threadIdx.x == 0) dot_part[blockIdx.x] = sblk;
    }
}

// ---- expansion kernels (upper BCOO -> both triangles, rows sorted by column)
__global__ void k_colrow_keys(const int* __restrict__ rows, const int* __restrict__ cols, int nnz, int kbits,
                              uint64_t* __restrict__ keys, int* __restrict__ idx)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    keys[k] = ((uint64_t)(uint32_t)cols[k] << kbits) | (uint64_t)(uint32_t)rows[k];
    idx[k] = k;
}

// up_ptr[i] = first upper block of row i; col_ptr[j] = first sorted entry of column j.
__global__ void k_ptrs(const int* __restrict__ rows, const uint64_t* __restrict__ skeys, int nnz, int n, int kbits,
                       int* __restrict__ up_ptr, int* __restrict__ col_ptr)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    int lo = 0, hi = nnz;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (rows[mid] < i) lo = mid + 1; else hi = mid;
    }
    up_ptr[i] = lo;
    const uint64_t key = (uint64_t)(uint32_t)i << kbits;
    lo = 0;
    hi = nnz;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (skeys[mid] < key) lo = mid + 1; else hi = mid;
    }
    col_ptr[i] = lo;
}

// strict lower count of row i = entries of column i with row < i
__global__ void k_row_len(const int* __restrict__ up_ptr, const int* __restrict__ col_ptr,
                          const uint64_t* __restrict__ skeys, int n, int kbits, int* __restrict__ low_cnt,
                          int* __restrict__ full_len)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int cb = col_ptr[i], ce = col_ptr[i + 1];
    int low = ce - cb;
    if (low > 0) {
        const uint64_t last = skeys[ce - 1];
        const int r = (int)(last & ((1ull << kbits) - 1ull));
        if (r == i) --low;  // the diagonal block closes its column
    }
    low_cnt[i] = low;
    full_len[i] = low + (up_ptr[i + 1] - up_ptr[i]);
}

__device__ __forceinline__ void copy9(const double* __restrict__ s, double* __restrict__ d, bool tr)
{
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) d[c * 3 + r] = tr ? s[r * 3 + c] : s[c * 3 + r];
}

__global__ void k_scatter_upper(const double* __restrict__ vals, const int* __restrict__ rows,
                                const int* __restrict__ cols, int nnz, const int* __restrict__ up_ptr,
                                const int* __restrict__ low_cnt, const int* __restrict__ full_ptr,
                                double* __restrict__ fv, int* __restrict__ fc)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const int i = rows[k];
    const int pos = full_ptr[i] + low_cnt[i] + (k - up_ptr[i]);
    copy9(vals + 9 * (long)k, fv + 9 * (long)pos, false);
    fc[pos] = cols[k];
}

__global__ void k_scatter_lower(const double* __restrict__ vals, const uint64_t* __restrict__ skeys,
                                const int* __restrict__ sidx, int nnz, int kbits, const int* __restrict__ col_ptr,
                                const int* __restrict__ full_ptr, double* __restrict__ fv, int* __restrict__ fc)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nnz) return;
    const uint64_t key = skeys[s];
    const int j = (int)(key >> kbits);
    const int k = (int)(key & ((1ull << kbits) - 1ull));
    if (k == j) return;  // diagonal block: placed by the upper scatter
    const int pos = full_ptr[j] + (s - col_ptr[j]);
    copy9(vals + 9 * (long)sidx[s], fv + 9 * (long)pos, true);
    fc[pos] = k;
}

// ---- SELL packing
__global__ void k_len_keys(const int* __restrict__ row_ptr, int n, uint32_t* __restrict__ key, int* __restrict__ iota)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int len = row_ptr[i + 1] - row_ptr[i];
    key[i] = 0xFFFFu - (uint32_t)(len < 0xFFFF ? len : 0xFFFF);  // ascending key = descending length
    iota[i] = i;
}

__global__ void k_slice_w(const int* __restrict__ perm, const int* __restrict__ row_ptr, int n, int n_slices,
                          int* __restrict__ slice_w)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > n_slices) return;
    if (q == n_slices) {
        slice_w[q] = 0;  // scan sentinel
        return;
    }
    const int i = perm[q * 64];  // longest row of the slice
    slice_w[q] = row_ptr[i + 1] - row_ptr[i];
}

template <bool kF32>
__global__ void k_sell_pack(const double* __restrict__ fv, const int* __restrict__ fc,
                            const int* __restrict__ row_ptr, const int* __restrict__ perm, int n,
                            const int* __restrict__ slice_ptr, const int* __restrict__ slice_w,
                            double* __restrict__ sv, float* __restrict__ sv32, int* __restrict__ sc)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int q = p >> 6, lane = p & 63;
    if (q * 64 >= ((n + 63) / 64) * 64) return;
    const int i = p < n ? perm[p] : -1;
    const int beg = i >= 0 ? row_ptr[i] : 0;
    const int len = i >= 0 ? row_ptr[i + 1] - beg : 0;
    const long S = slice_ptr[q];
    const int W = slice_w[q];
    for (int s = 0; s < W; ++s) {
        const long slot = S + s;
        const bool real = s < len;
        sc[slot * 64 + lane] = real ? fc[beg + s] : (i >= 0 ? i : 0);
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            const double v = real ? fv[9 * (long)(beg + s) + c] : 0.0;
            sv[(slot * 9 + c) * 64 + lane] = v;
            if (kF32) sv32[(slot * 9 + c) * 64 + lane] = (float)v;
        }
    }
}

inline int g256(long n) { return (int)((n + 255) / 256); }

int grid_sell(int n_slices)
{
    int g = (n_slices + kWaves - 1) / kWaves;
    if (g > kMaxGrid) g = kMaxGrid;
    if (g < 1) g = 1;
    return g;
}

}  // namespace

void FullMatrix::build(hipStream_t s, const double* vals, const int* rows, const int* cols, int nnz_up, int n,
                       bool want_f32)
{
    SGIPC_CHECK(nnz_up > 0 && n > 0, SGIPC_EMPTY, "empty matrix");
    n_rows = n;
    has_f32 = want_f32;
    int kbits = 1;
    while (kbits < 31 && (1L << kbits) < (long)n) ++kbits;
    keys_.reserve(nnz_up);
    skeys_.reserve(nnz_up);
    idx_.reserve(nnz_up);
    sidx_.reserve(nnz_up);
    up_ptr_.reserve(n + 1);
    col_ptr_.reserve(n + 1);
    low_cnt_.reserve(n);
    full_len_.reserve(n + 1);
    row_ptr_.reserve(n + 1);
    len_key_.reserve(n);
    len_key_sorted_.reserve(n);
    iota_.reserve(n);
    perm.reserve(n);
    n_slices = (n + 63) / 64;
    slice_w.reserve(n_slices + 1);
    slice_ptr.reserve(n_slices + 1);

    // ---- expansion
    k_colrow_keys<<<g256(nnz_up), 256, 0, s>>>(rows, cols, nnz_up, kbits, keys_.get(), idx_.get());
    SGIPC_LAUNCH_CHECK();
    size_t need = 0, need2 = 0, need3 = 0, need4 = 0;
    SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, need, keys_.get(), skeys_.get(), idx_.get(), sidx_.get(),
                                        (size_t)nnz_up, 0, 2 * kbits, s));
    SGIPC_HIP(rocprim::exclusive_scan(nullptr, need2, full_len_.get(), row_ptr_.get(), 0, (size_t)n + 1,
                                      rocprim::plus<int>(), s));
    SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, need3, reinterpret_cast<uint32_t*>(len_key_.get()),
                                        reinterpret_cast<uint32_t*>(len_key_sorted_.get()), iota_.get(), perm.get(),
                                        (size_t)n, 0, 16, s));
    SGIPC_HIP(rocprim::exclusive_scan(nullptr, need4, slice_w.get(), slice_ptr.get(), 0, (size_t)n_slices + 1,
                                      rocprim::plus<int>(), s));
    size_t nt = need;
    if (need2 > nt) nt = need2;
    if (need3 > nt) nt = need3;
    if (need4 > nt) nt = need4;
    temp_.reserve(nt);
    SGIPC_HIP(rocprim::radix_sort_pairs(temp_.get(), need, keys_.get(), skeys_.get(), idx_.get(), sidx_.get(),
                                        (size_t)nnz_up, 0, 2 * kbits, s));
    k_ptrs<<<g256(n + 1), 256, 0, s>>>(rows, skeys_.get(), nnz_up, n, kbits, up_ptr_.get(), col_ptr_.get());
    SGIPC_LAUNCH_CHECK();
    k_row_len<<<g256(n), 256, 0, s>>>(up_ptr_.get(), col_ptr_.get(), skeys_.get(), n, kbits, low_cnt_.get(),
                                      full_len_.get());
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(hipMemsetAsync(full_len_.get() + n, 0, sizeof(int), s));
    SGIPC_HIP(rocprim::exclusive_scan(temp_.get(), need2, full_len_.get(), row_ptr_.get(), 0, (size_t)n + 1,
                                      rocprim::plus<int>(), s));
    int h_nnz = 0;
    SGIPC_HIP(hipMemcpyAsync(&h_nnz, row_ptr_.get() + n, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    nnz = h_nnz;
    fvals_.reserve((size_t)nnz * 9);
    fcols_.reserve(nnz);
    k_scatter_upper<<<g256(nnz_up), 256, 0, s>>>(vals, rows, cols, nnz_up, up_ptr_.get(), low_cnt_.get(),
                                                 row_ptr_.get(), fvals_.get(), fcols_.get());
    SGIPC_LAUNCH_CHECK();
    k_scatter_lower<<<g256(nnz_up), 256, 0, s>>>(vals, skeys_.get(), sidx_.get(), nnz_up, kbits, col_ptr_.get(),
                                                 row_ptr_.get(), fvals_.get(), fcols_.get());
    SGIPC_LAUNCH_CHECK();

    // ---- SELL-64-sigma packing
    k_len_keys<<<g256(n), 256, 0, s>>>(row_ptr_.get(), n, reinterpret_cast<uint32_t*>(len_key_.get()), iota_.get());
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(rocprim::radix_sort_pairs(temp_.get(), need3, reinterpret_cast<uint32_t*>(len_key_.get()),
                                        reinterpret_cast<uint32_t*>(len_key_sorted_.get()), iota_.get(), perm.get(),
                                        (size_t)n, 0, 16, s));
    k_slice_w<<<g256(n_slices + 1), 256, 0, s>>>(perm.get(), row_ptr_.get(), n, n_slices, slice_w.get());
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(rocprim::exclusive_scan(temp_.get(), need4, slice_w.get(), slice_ptr.get(), 0, (size_t)n_slices + 1,
                                      rocprim::plus<int>(), s));
    int h_slots = 0;
    SGIPC_HIP(hipMemcpyAsync(&h_slots, slice_ptr.get() + n_slices, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    n_slots = h_slots;
    sell_vals.reserve((size_t)n_slots * 64 * 9);
    sell_cols.reserve((size_t)n_slots * 64);
    if (want_f32) sell_vals32.reserve((size_t)n_slots * 64 * 9);
    const long threads = (long)n_slices * 64;
    if (want_f32)
        k_sell_pack<true><<<g256(threads), 256, 0, s>>>(fvals_.get(), fcols_.get(), row_ptr_.get(), perm.get(), n,
                                                        slice_ptr.get(), slice_w.get(), sell_vals.get(),
                                                        sell_vals32.get(), sell_cols.get());
    else
        k_sell_pack<false><<<g256(threads), 256, 0, s>>>(fvals_.get(), fcols_.get(), row_ptr_.get(), perm.get(), n,
                                                         slice_ptr.get(), slice_w.get(), sell_vals.get(), nullptr,
                                                         sell_cols.get());
    SGIPC_LAUNCH_CHECK();
}

int FullMatrix::partials() const { return grid_sell(n_slices); }

template <typename TV>
int FullMatrix::apply(hipStream_t s, TV a, const TV* x, TV b, TV* y, double* dot_part, const int* done) const
{
    const int g = grid_sell(n_slices);
    if constexpr (sizeof(TV) == 8) {
        if (dot_part)
            k_spmv_sell<double, double, true><<<g, kThreads, 0, s>>>(sell_vals.get(), sell_cols.get(), slice_ptr.get(),
                                                                    slice_w.get(), perm.get(), n_rows, n_slices, a, x,
                                                                    b, y, dot_part, done);
        else
            k_spmv_sell<double, double, false><<<g, kThreads, 0, s>>>(sell_vals.get(), sell_cols.get(),
                                                                     slice_ptr.get(), slice_w.get(), perm.get(), n_rows,
                                                                     n_slices, a, x, b, y, nullptr, done);
    } else {
        SGIPC_CHECK(has_f32, SGIPC_BAD_ARG, "FP32 copy of the matrix was not built");
        if (dot_part)
            k_spmv_sell<float, float, true><<<g, kThreads, 0, s>>>(sell_vals32.get(), sell_cols.get(), slice_ptr.get(),
                                                                  slice_w.get(), perm.get(), n_rows, n_slices, a, x, b,
                                                                  y, dot_part, done);
        else
            k_spmv_sell<float, float, false><<<g, kThreads, 0, s>>>(sell_vals32.get(), sell_cols.get(),
                                                                   slice_ptr.get(), slice_w.get(), perm.get(), n_rows,
                                                                   n_slices, a, x, b, y, nullptr, done);
    }
    SGIPC_LAUNCH_CHECK();
    return g;
}
template int FullMatrix::apply<double>(hipStream_t, double, const double*, double, double*, double*, const int*) const;
template int FullMatrix::apply<float>(hipStream_t, float, const float*, float, float*, double*, const int*) const;

}  // namespace sgipc
