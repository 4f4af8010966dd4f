// Symmetric upper-BCOO 3x3 SpMV for gfx950 ("SRBK": segmented row reduction
// plus transposed scatter), y = a*A*x + b*y with A = U + U^T - diag(U).
//
// Follows the semantics of Spmv::srbk_spmv (StiffGIPC/linear_system/utils/
// spmv.cu:188-313): for every stored block (i,j,H), i <= j, row i receives
// a*H*x_j and, when i != j, row j receives a*H^T*x_i.
//
// MI355X mapping: one wavefront streams a chunk of 64 consecutive blocks.  The
// chunk's 64x72 bytes of block values are fetched with fully coalesced 16-byte
// loads into LDS and re-read per lane, so HBM sees one contiguous stream; the
// row part is reduced across the wave with a segmented scan keyed on the row
// index (rows are sorted) and committed with one atomic per row segment; the
// transposed part is committed with hardware FP atomics.  The grid is sized to
// keep every CU busy (grid-stride over chunks) and optionally emits one partial
// of x^T A x per workgroup (the PCG p.Ap product) so no extra pass is needed.
#include "pcg.hpp"
#include "sgipc_internal.hpp"
#include "wave_ops.hpp"

namespace sgipc {

namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kThreads = kWavesPerBlock * 64;

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
struct Stage;
template <>
struct Stage<double> {
    using V = double2;  // 16-byte vector
    static constexpr int kPerVec = 2;
};
template <>
struct Stage<float> {
    using V = float4;
    static constexpr int kPerVec = 4;
};

template <typename T, bool kDot, bool kAligned>
__global__ __launch_bounds__(kThreads) void spmv_srbk_kernel(
    T a, const T* __restrict__ vals, const int* __restrict__ rows, const int* __restrict__ cols,
    int nnz, const T* __restrict__ x, T* __restrict__ y, double* __restrict__ partials,
    const int* __restrict__ done)
{
    if (done != nullptr && *done) return;
    using V = typename Stage<T>::V;
    constexpr int kPV = Stage<T>::kPerVec;
    constexpr int kChunkVals = 64 * 9;
    __shared__ __align__(16) T stage[kWavesPerBlock][kChunkVals];
    __shared__ double red[16];

    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const long nchunks = ((long)nnz + 63) / 64;
    double dot_acc = 0.0;
    T* sv = stage[w];

    for (long c = (long)blockIdx.x * kWavesPerBlock + w; c < nchunks;
         c += (long)gridDim.x * kWavesPerBlock) {
        const long base = c * 64;
        const int cnt = (int)min(64L, (long)nnz - base);
        // coalesced staging of cnt*9 values (the chunk start is 16-byte aligned)
        const int nv = cnt * 9;
        const int nvec = kAligned ? nv / kPV : 0;
        if (kAligned) {
            const V* src = reinterpret_cast<const V*>(vals + base * 9);
            for (int e = lane; e < nvec; e += 64) reinterpret_cast<V*>(sv)[e] = src[e];
        }
        for (int e = nvec * kPV + lane; e < nv; e += 64) sv[e] = vals[base * 9 + e];
        wave_lds_sync();

        const bool valid = lane < cnt;
        const long k = base + lane;
        int i = valid ? rows[k] : -1 - lane;
        int j = valid ? cols[k] : 0;
        T v0 = 0, v1 = 0, v2 = 0;
        if (valid) {
            const T* H = sv + lane * 9;  // column-major 3x3
            const T xj0 = x[3 * (long)j], xj1 = x[3 * (long)j + 1], xj2 = x[3 * (long)j + 2];
            v0 = H[0] * xj0 + H[3] * xj1 + H[6] * xj2;
            v1 = H[1] * xj0 + H[4] * xj1 + H[7] * xj2;
            v2 = H[2] * xj0 + H[5] * xj1 + H[8] * xj2;
            const T xi0 = x[3 * (long)i], xi1 = x[3 * (long)i + 1], xi2 = x[3 * (long)i + 2];
            if (kDot) {
                double d = (double)xi0 * v0 + (double)xi1 * v1 + (double)xi2 * v2;
                dot_acc += (i != j) ? 2.0 * d : d;
            }
            if (i != j) {
                const T u0 = H[0] * xi0 + H[1] * xi1 + H[2] * xi2;
                const T u1 = H[3] * xi0 + H[4] * xi1 + H[5] * xi2;
                const T u2 = H[6] * xi0 + H[7] * xi1 + H[8] * xi2;
                atomic_add(y + 3 * (long)j, a * u0);
                atomic_add(y + 3 * (long)j + 1, a * u1);
                atomic_add(y + 3 * (long)j + 2, a * u2);
            }
        }
        wave_lds_sync();  // stage buffer is rewritten by the next chunk

        // segmented reduction of the row part, keyed on the sorted row index
        const int prev = __shfl_up(i, 1, 64);
        const int next = __shfl_down(i, 1, 64);
        const bool head = (lane == 0) || (prev != i);
        const bool tail = (lane == 63) || (next != i);
        const uint64_t heads = __ballot(head);
        const int start = seg_start_from_heads(heads, lane);
        v0 = seg_inclusive_sum(v0, lane, start);
        v1 = seg_inclusive_sum(v1, lane, start);
        v2 = seg_inclusive_sum(v2, lane, start);
        if (valid && tail) {
            atomic_add(y + 3 * (long)i, a * v0);
            atomic_add(y + 3 * (long)i + 1, a * v1);
            atomic_add(y + 3 * (long)i + 2, a * v2);
        }
    }
    if (kDot) {
        double s = block_sum(dot_acc, red);
        if (threadIdx.x == 0) partials[blockIdx.x] = s;
    }
}

template <typename T>
__global__ void scale_kernel(T* __restrict__ y, T b, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (b == T(0)) ? T(0) : b * y[i];
}

__global__ void f64_to_f32_kernel(const double* __restrict__ in, float* __restrict__ out, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

int spmv_grid(int nnz)
{
    long chunks = ((long)nnz + 63) / 64;
    long g = (chunks + kWavesPerBlock - 1) / kWavesPerBlock;
    if (g > 2048) g = 2048;
    if (g < 1) g = 1;
    return (int)g;
}

}  // namespace

int spmv_partials_needed(int nnz) { return spmv_grid(nnz); }

void launch_spmv_f64(hipStream_t s, double a, const double* vals, const int* rows,
                     const int* cols, int nnz, const double* x, double b, double* y, int n_rows,
                     double* pap_partials, int* n_partials_out)
{
    long n3 = 3L * n_rows;
    if (n3 > 0 && b != 1.0) {
        scale_kernel<double><<<(unsigned)((n3 + 255) / 256), 256, 0, s>>>(y, b, n3);
        SGIPC_LAUNCH_CHECK();
    }
    int g = spmv_grid(nnz);
    if (n_partials_out) *n_partials_out = g;
    if (nnz <= 0) return;
    const bool al = (reinterpret_cast<uintptr_t>(vals) % 16) == 0;
    if (pap_partials) {
        if (al) spmv_srbk_kernel<double, true, true><<<g, kThreads, 0, s>>>(a, vals, rows, cols, nnz, x, y, pap_partials, nullptr);
        else spmv_srbk_kernel<double, true, false><<<g, kThreads, 0, s>>>(a, vals, rows, cols, nnz, x, y, pap_partials, nullptr);
    } else {
        if (al) spmv_srbk_kernel<double, false, true><<<g, kThreads, 0, s>>>(a, vals, rows, cols, nnz, x, y, nullptr, nullptr);
        else spmv_srbk_kernel<double, false, false><<<g, kThreads, 0, s>>>(a, vals, rows, cols, nnz, x, y, nullptr, nullptr);
    }
    SGIPC_LAUNCH_CHECK();
}

void launch_spmv_f32(hipStream_t s, float a, const float* vals, const int* rows, const int* cols,
                     int nnz, const float* x, float b, float* y, int n_rows)
{
    long n3 = 3L * n_rows;
    if (n3 > 0 && b != 1.0f) {
        scale_kernel<float><<<(unsigned)((n3 + 255) / 256), 256, 0, s>>>(y, b, n3);
        SGIPC_LAUNCH_CHECK();
    }
    if (nnz <= 0) return;
    const bool al = (reinterpret_cast<uintptr_t>(vals) % 16) == 0;
    if (al) spmv_srbk_kernel<float, false, true><<<spmv_grid(nnz), kThreads, 0, s>>>(a, vals, rows, cols, nnz, x, y, nullptr, nullptr);
    else spmv_srbk_kernel<float, false, false><<<spmv_grid(nnz), kThreads, 0, s>>>(a, vals, rows, cols, nnz, x, y, nullptr, nullptr);
    SGIPC_LAUNCH_CHECK();
}

template <typename T>
void launch_spmv_pcg(hipStream_t s, const T* vals, const int* rows, const int* cols, int nnz,
                     const T* p, T* Ap, double* pap_part, const PcgCtl* ctl, int* n_part)
{
    int g = spmv_grid(nnz);
    *n_part = g;
    const int* done = ctl ? &ctl->done : nullptr;
    const bool al = (reinterpret_cast<uintptr_t>(vals) % 16) == 0;
    if (al)
        spmv_srbk_kernel<T, true, true><<<g, kThreads, 0, s>>>(T(1), vals, rows, cols, nnz, p, Ap, pap_part, done);
    else
        spmv_srbk_kernel<T, true, false><<<g, kThreads, 0, s>>>(T(1), vals, rows, cols, nnz, p, Ap, pap_part, done);
    SGIPC_LAUNCH_CHECK();
}

void launch_spmv_pcg_f64(hipStream_t s, const double* vals, const int* rows, const int* cols,
                         int nnz, const double* p, double* Ap, double* pap_part, const PcgCtl* ctl,
                         int* n_part)
{
    launch_spmv_pcg<double>(s, vals, rows, cols, nnz, p, Ap, pap_part, ctl, n_part);
}

void launch_spmv_pcg_f32(hipStream_t s, const float* vals, const int* rows, const int* cols,
                         int nnz, const float* p, float* Ap, double* pap_part, const PcgCtl* ctl,
                         int* n_part)
{
    launch_spmv_pcg<float>(s, vals, rows, cols, nnz, p, Ap, pap_part, ctl, n_part);
}

void launch_f64_to_f32(hipStream_t s, const double* in, float* out, size_t n)
{
    if (!n) return;
    f64_to_f32_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(in, out, (long)n);
    SGIPC_LAUNCH_CHECK();
}

}  // namespace sgipc
