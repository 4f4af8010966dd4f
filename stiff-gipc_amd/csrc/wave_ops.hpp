// Wavefront (64-lane) primitives for gfx950: reductions, segmented scans and
// lane masks built on ds_bpermute / DPP-backed shuffles.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sgipc {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        T o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// Index of the highest set bit of m that is <= lane (m must have one such bit).
__device__ __forceinline__ int seg_start_from_heads(uint64_t heads, int lane)
{
    uint64_t upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    return 63 - __clzll(heads & upto);
}

// Inclusive segmented prefix sum across the wave: after the call the last lane
// of every segment holds the segment total.  `start` is the first lane of the
// caller's segment.
template <typename T>
__device__ __forceinline__ T seg_inclusive_sum(T v, int lane, int start)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        T t = __shfl_up(v, off, 64);
        if (lane - off >= start) v += t;
    }
    return v;
}

// Block-wide deterministic sum of `v` (blockDim.x a multiple of 64, <= 1024).
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* smem /* >= 16 */)
{
    v = wave_sum(v);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) smem[w] = v;
    __syncthreads();
    int nw = blockDim.x >> 6;
    T s = 0;
    if (threadIdx.x < 64) {
        s = (int)threadIdx.x < nw ? smem[threadIdx.x] : T(0);
        s = wave_sum(s);
    }
    __syncthreads();
    if (threadIdx.x == 0) smem[0] = s;
    __syncthreads();
    T r = smem[0];
    __syncthreads();
    return r;
}

// Deterministic sum of n partials by a whole workgroup (fixed order).
__device__ __forceinline__ double block_sum_array(const double* __restrict__ a, int n, double* smem)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += a[i];
    return block_sum(s, smem);
}

__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }

}  // namespace sgipc
