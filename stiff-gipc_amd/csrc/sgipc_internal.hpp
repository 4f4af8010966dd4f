// Internal declarations of the MI355X (gfx950) Newton linear-solve library.
// Host runtime objects (context, device buffers, error state) and the kernel
// launchers shared by the translation units.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sgipc.h"

namespace sgipc {

// A failed HIP call or a rejected argument; converted to a status code at the
// C ABI and never allowed to cross it.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define SGIPC_HIP(expr)                                                                  \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess)                                                            \
            throw ::sgipc::Error(SGIPC_HIP_ERROR, std::string(#expr) + ": " +           \
                                                      hipGetErrorString(_e));            \
    } while (0)

#define SGIPC_CHECK(cond, code, msg)                                                     \
    do {                                                                                 \
        if (!(cond)) throw ::sgipc::Error((code), (msg));                               \
    } while (0)

#define SGIPC_LAUNCH_CHECK() SGIPC_HIP(hipGetLastError())

constexpr int kWave = 64;

// Owned device allocation.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // Grow to at least `count` elements (contents not preserved).
    void reserve(size_t count)
    {
        if (count <= n && p) return;
        release();
        size_t bytes = (count ? count : 1) * sizeof(T);
        SGIPC_HIP(hipMalloc(reinterpret_cast<void**>(&p), bytes));
        n = count;
    }
    T* get() const { return p; }
};

// A second stream (default priority) for build work that is independent of
// what the main stream runs meanwhile.  begin(main): work queued on the side
// stream from now on starts after everything queued on main so far; end(main):
// work queued on main from now on starts after everything queued on the side
// stream so far.  Made on first use with the owner's device current.
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    SideStream() = default;
    SideStream(const SideStream&) = delete;
    SideStream& operator=(const SideStream&) = delete;
    ~SideStream()
    {
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (s) (void)hipStreamDestroy(s);
    }
    void begin(hipStream_t main)
    {
        if (!s) {
            SGIPC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            SGIPC_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
            SGIPC_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming));
        }
        SGIPC_HIP(hipEventRecord(fork, main));
        SGIPC_HIP(hipStreamWaitEvent(s, fork, 0));
    }
    void end(hipStream_t main)
    {
        SGIPC_HIP(hipEventRecord(join, s));
        SGIPC_HIP(hipStreamWaitEvent(main, join, 0));
    }
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string last_error;
    int num_cus = 256;
    // scratch for reductions (per-workgroup partials) and scalar slots
    DevBuf<double> red;       // partial sums
    DevBuf<int> flags;        // done flag / counters
    double* pinned = nullptr; // host-pinned scalars
    void activate() const { SGIPC_HIP(hipSetDevice(device)); }
};

// ----------------------------------------------------------------- SpMV
void launch_spmv_f64(hipStream_t s, double a, const double* vals, const int* rows,
                     const int* cols, int nnz, const double* x, double b, double* y, int n_rows,
                     double* pap_partials, int* n_partials_out);
void launch_spmv_f32(hipStream_t s, float a, const float* vals, const int* rows, const int* cols,
                     int nnz, const float* x, float b, float* y, int n_rows);
int spmv_partials_needed(int nnz);

// ----------------------------------------------------------------- conversion helpers
void launch_f64_to_f32(hipStream_t s, const double* in, float* out, size_t n);

}  // namespace sgipc
