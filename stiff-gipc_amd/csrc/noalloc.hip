// Allocation-free triplet conversion: sgipc_convert_srbk's sort + reduce-by-key
// run in a fixed device array instead of buffers the library allocates, for
// callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation).
//
// The array `scratch` below is a file-scope __device__ array: each GPU the
// library runs on holds its own copy (kScratchBytes) until the process ends.
// In host code its name is only a host-side stand-in, so the host never
// passes it to a kernel: noalloc_setup resolves its device address once per
// device (hipGetSymbolAddress, with that device current) and every launch
// and copy uses that address.  Every call on one GPU shares that GPU's
// array, so calls that overlap on one GPU (two streams, two threads) race;
// each call clears what must start at zero itself.
#include <mutex>

#include "assembly.hpp"
#include "capi_internal.hpp"

namespace sgipc {

namespace {

// Largest triplet count one call converts (C0's 16^3 grid has 206,596).
constexpr long kMaxTriplets = 1L << 18;
// keys, sorted keys (8 B each), indices, sorted indices, flags, scan (4 B
// each) and block starts (n + 1) take 36 B per triplet; the rest holds the
// scratch ints and the radix sort / scan temporary storage.
constexpr size_t kScratchBytes = 24u << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
unsigned char* g_base[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Device address of `scratch` on `device`, which must be current.
unsigned char* scratch_base(int device)
{
    SGIPC_CHECK(device >= 0 && device < kMaxDevices, SGIPC_BAD_ARG, "device index beyond the allocation-free table");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_done[device]) {
        void* p = nullptr;
        g_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_base[device] = g_err[device] == hipSuccess ? static_cast<unsigned char*>(p) : nullptr;
        g_done[device] = true;
    }
    SGIPC_CHECK(g_err[device] == hipSuccess && g_base[device] != nullptr, SGIPC_HIP_ERROR,
                std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(g_err[device]));
    return g_base[device];
}

}  // namespace

void noalloc_setup(int device)
{
    try {
        (void)scratch_base(device);
    } catch (const Error&) {
        // reported by the entry point's first call
    }
}

}  // namespace sgipc

using namespace sgipc;

extern "C" int sgipc_convert_srbk_noalloc(sgipc_ctx* ctx, const double* tvals, const int* trows, const int* tcols,
                                          int n_triplets, double* ovals, int* orows, int* ocols, int capacity,
                                          int* out_nnz)
{
    if (!ctx || !out_nnz || n_triplets < 0 || capacity < 0) return SGIPC_BAD_ARG;
    *out_nnz = 0;
    if (n_triplets == 0) return SGIPC_EMPTY;
    if (!tvals || !trows || !tcols || !ovals || !orows || !ocols) return SGIPC_BAD_ARG;
    int unique = 0;
    const int st = guarded(ctx, [&] {
        SGIPC_CHECK(n_triplets <= kMaxTriplets, SGIPC_CAPACITY,
                    "n_triplets exceeds the allocation-free maximum of 262144");
        unsigned char* base = scratch_base(ctx->c.device);
        const long n = n_triplets;
        ConvertScratch w;
        size_t off = 0;
        auto take = [&](size_t bytes) {
            unsigned char* p = base + off;
            off += align256(bytes);
            return p;
        };
        w.keys = reinterpret_cast<uint64_t*>(take(sizeof(uint64_t) * n));
        w.skeys = reinterpret_cast<uint64_t*>(take(sizeof(uint64_t) * n));
        w.idx = reinterpret_cast<int*>(take(sizeof(int) * n));
        w.sidx = reinterpret_cast<int*>(take(sizeof(int) * n));
        w.flag = reinterpret_cast<int*>(take(sizeof(int) * n));
        w.pos = reinterpret_cast<int*>(take(sizeof(int) * n));
        w.head = reinterpret_cast<int*>(take(sizeof(int) * (n + 1)));
        w.ints = reinterpret_cast<int*>(take(sizeof(int) * 4));
        SGIPC_CHECK(off < kScratchBytes, SGIPC_CAPACITY, "allocation-free scratch too small");
        w.temp = base + off;
        w.temp_bytes = kScratchBytes - off;
        hipStream_t s = ctx->c.stream;
        const int mx = convert_max_index(s, trows, tcols, n, w.ints);
        convert_into(s, w, tvals, trows, tcols, n, mx + 1, ovals, orows, ocols, capacity, &unique);
    });
    // a capacity failure reports the unique count the output needs
    *out_nnz = unique;
    return st;
}
