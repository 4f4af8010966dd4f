// Slot remapping kernels (memory-bound, one thread per triplet / pair / vertex).
#include "slots.hpp"

namespace sgipc {

namespace {

constexpr int kT = 256;

inline int blocks(long n) { return (int)((n + kT - 1) / kT); }

__global__ __launch_bounds__(kT) void k_slot_triplets(long n, int n_verts, const int* __restrict__ slot_of,
                                                      double* __restrict__ v, int* __restrict__ rows,
                                                      int* __restrict__ cols)
{
    const long i = (long)blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const int r = rows[i], c = cols[i];
    if (r < 0 || c < 0 || r >= n_verts || c >= n_verts) {
        rows[i] = -1;  // the converter's index check reports it
        return;
    }
    const int a = slot_of[r], b = slot_of[c];
    if (a <= b) {
        rows[i] = a;
        cols[i] = b;
        return;
    }
    rows[i] = b;
    cols[i] = a;
    double* h = v + 9 * i;  // column-major 3x3: h[c*3 + r]
    double t;
    t = h[1]; h[1] = h[3]; h[3] = t;
    t = h[2]; h[2] = h[6]; h[6] = t;
    t = h[5]; h[5] = h[7]; h[7] = t;
}

// a vertex id, or its -v-1 encoding
__device__ __forceinline__ int rename(int e, const int* __restrict__ slot_of)
{
    return e >= 0 ? slot_of[e] : -slot_of[-e - 1] - 1;
}

// The encodings (StiffGIPC/mlbvh.cu:539-760, GIPC.cu:3149-4986):
// EE (a0, a1, b0, b1|-b1-1); PT (-p-1, t0, t1, t2); PE (-p-1, a, b, w<0);
// PP (-p-1, q, -1, w); nearly parallel EE as PE/PP shapes with every vertex
// entry encoded (-v-1).
__global__ __launch_bounds__(kT) void k_slot_pairs(int n, const int* __restrict__ slot_of,
                                                   const int4* __restrict__ in, int4* __restrict__ out)
{
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const int4 p = in[i];
    int4 o = p;
    if (p.x >= 0) {
        o.x = slot_of[p.x];
        o.y = slot_of[p.y];
        o.z = slot_of[p.z];
        o.w = rename(p.w, slot_of);
    } else {
        o.x = rename(p.x, slot_of);
        if (p.z < 0) {
            if (p.y < 0) {
                o.y = rename(p.y, slot_of);
                o.z = rename(p.z, slot_of);
                o.w = rename(p.w, slot_of);
            } else {
                o.y = slot_of[p.y];  // point-point: z = -1 and w are not vertices
            }
        } else if (p.w < 0) {
            if (p.y < 0) {
                o.y = rename(p.y, slot_of);
                o.z = slot_of[p.z];
                o.w = rename(p.w, slot_of);
            } else {
                o.y = slot_of[p.y];  // point-edge: w is not a vertex
                o.z = slot_of[p.z];
            }
        } else {
            o.y = slot_of[p.y];
            o.z = slot_of[p.z];
            o.w = slot_of[p.w];
        }
    }
    out[i] = o;
}

__global__ __launch_bounds__(kT) void k_zero(long n, double* __restrict__ a)
{
    const long i = (long)blockIdx.x * kT + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

__global__ __launch_bounds__(kT) void k_slot_scatter(int n, const int* __restrict__ slot_of,
                                                     const double* __restrict__ v, double* __restrict__ out)
{
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const long s = 3L * slot_of[i];
    out[s] = v[3L * i];
    out[s + 1] = v[3L * i + 1];
    out[s + 2] = v[3L * i + 2];
}

__global__ __launch_bounds__(kT) void k_slot_gather(int n, const int* __restrict__ slot_of,
                                                    const double* __restrict__ in, double* __restrict__ v)
{
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const long s = 3L * slot_of[i];
    v[3L * i] = in[s];
    v[3L * i + 1] = in[s + 1];
    v[3L * i + 2] = in[s + 2];
}

}  // namespace

void launch_slot_triplets(hipStream_t s, long n, int n_verts, const int* slot_of, double* vals, int* rows, int* cols)
{
    if (n <= 0) return;
    k_slot_triplets<<<blocks(n), kT, 0, s>>>(n, n_verts, slot_of, vals, rows, cols);
    SGIPC_LAUNCH_CHECK();
}

void launch_slot_pairs(hipStream_t s, int n, const int* slot_of, const int* pairs, int* out)
{
    if (n <= 0) return;
    k_slot_pairs<<<blocks(n), kT, 0, s>>>(n, slot_of, reinterpret_cast<const int4*>(pairs),
                                          reinterpret_cast<int4*>(out));
    SGIPC_LAUNCH_CHECK();
}

void launch_slot_scatter(hipStream_t s, int n, int n_slots, const int* slot_of, const double* v, double* out)
{
    k_zero<<<blocks(3L * n_slots), kT, 0, s>>>(3L * n_slots, out);
    SGIPC_LAUNCH_CHECK();
    if (n <= 0) return;
    k_slot_scatter<<<blocks(n), kT, 0, s>>>(n, slot_of, v, out);
    SGIPC_LAUNCH_CHECK();
}

void launch_slot_gather(hipStream_t s, int n, const int* slot_of, const double* in, double* v)
{
    if (n <= 0) return;
    k_slot_gather<<<blocks(n), kT, 0, s>>>(n, slot_of, in, v);
    SGIPC_LAUNCH_CHECK();
}

}  // namespace sgipc
