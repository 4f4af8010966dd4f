// Learned coarse level on gfx950: level-1 graph build, BF16-MFMA layer
// kernels (inference and training), backward kernels and Adam.
//
// Layer kernel (one workgroup = 32 nodes = 96 rows, 4 waves):
//   1. gather: every lane owns one channel; a wave walks 8 nodes and, for
//      each, accumulates (Ahat V)_i = sum_j Ahat_ij V_j over the node's CSR
//      row (the block is a broadcast load, V_j is a coalesced 256-byte row);
//      U = [V, Ahat V] goes to LDS in bf16, V itself in float.
//   2. MFMA: Y (96 x 64) = U (96 x 128) W (128 x 64) with
//      v_mfma_f32_16x16x32_bf16; wave w owns output columns 16w..16w+15.
//   3. epilogue: V + relu(Y) (float), the decoder dot with w_out on the last
//      layer, and in training U (bf16) and Y (float) for the backward pass.
// All reductions are in fixed order: results do not depend on scheduling.
#include "mpnn.hpp"
#include "pcg.hpp"
#include "wave_ops.hpp"

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

namespace sgipc {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int C = kMpnnC;
constexpr int kRows = 3 * kMpnnTile;  // rows per layer workgroup
constexpr int kUld = 2 * C + 8;       // LDS row stride (bf16) of U
constexpr int kWld = 2 * C + 8;       // LDS row stride (bf16) of W^T
constexpr int kDld = C + 8;           // LDS row stride (bf16) of dY
constexpr int kDwRows = 384;          // rows per workgroup of the dW reduction
constexpr int kVec = 256;

// float -> bfloat16, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p)
{
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(p));
}

// ------------------------------------------------------------- graph build
// Directed entries of the symmetric operator from its upper blocks: entry
// (r, c) from upper block k, and (c, r) transposed when r != c.  Unused
// slots (diagonal blocks have one direction) get the largest key.
__global__ void k_g_keys(const int* __restrict__ rows, const int* __restrict__ cols, int nb,
                         uint64_t* __restrict__ keys, int* __restrict__ idx)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    const uint32_t r = (uint32_t)rows[k], c = (uint32_t)cols[k];
    keys[2 * k] = ((uint64_t)r << 32) | c;
    idx[2 * k] = k;
    if (r != c) {
        keys[2 * k + 1] = ((uint64_t)c << 32) | r;
        idx[2 * k + 1] = k | (int)0x80000000u;
    } else {
        keys[2 * k + 1] = ~0ull;
        idx[2 * k + 1] = -1;
    }
}

__global__ void k_g_count_diag(const int* __restrict__ rows, const int* __restrict__ cols, int nb, int* __restrict__ cnt)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool d = k < nb && rows[k] == cols[k];
    const uint64_t m = __ballot(d);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(cnt, __popcll(m));
}

// row_ptr from the sorted keys; entries past E hold ~0.
__global__ void k_g_rowptr(const uint64_t* __restrict__ sk, int E, int n, int* __restrict__ row_ptr)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e > E) return;
    const long prev = e == 0 ? -1 : (long)(sk[e - 1] >> 32);
    const long cur = e == E ? n : (long)(sk[e] >> 32);
    for (long r = prev + 1; r <= cur; ++r) row_ptr[r] = e;
}

__global__ void k_g_scale(const uint64_t* __restrict__ sk, const int* __restrict__ sidx, int E,
                          const double* __restrict__ vals, float* __restrict__ s)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const uint32_t r = (uint32_t)(sk[e] >> 32), c = (uint32_t)sk[e];
    if (r != c) return;
    const double* v = vals + 9L * (sidx[e] & 0x7FFFFFFF);
    const double d = (v[0] + v[4] + v[8]) / 3.0;
    s[r] = d > 0.0 ? (float)(1.0 / sqrt(d)) : 0.0f;
}

__global__ void k_g_blocks(const uint64_t* __restrict__ sk, const int* __restrict__ sidx, int E,
                           const double* __restrict__ vals, const float* __restrict__ s, int* __restrict__ col,
                           float* __restrict__ blk)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const uint32_t r = (uint32_t)(sk[e] >> 32), c = (uint32_t)sk[e];
    const int id = sidx[e];
    const bool tr = id < 0;
    const double* v = vals + 9L * (id & 0x7FFFFFFF);
    const double sc = (double)s[r] * (double)s[c];
    col[e] = (int)c;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            // upper block stored column-major: M[a][b] = v[3b + a]
            const double m = tr ? v[3 * a + b] : v[3 * b + a];
            blk[9L * e + 3 * a + b] = (float)(sc * m);
        }
}

// ------------------------------------------------------------------ layers
// Y = Ahat V on 64 channels: one wave per node (lanes = channels), so a
// launch keeps thousands of independent neighbour walks in flight.
__global__ __launch_bounds__(256) void k_spmm64(int n, const int* __restrict__ rp, const int* __restrict__ colx,
                                                const float* __restrict__ blk, const float* __restrict__ V,
                                                float* __restrict__ Y)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int ch = threadIdx.x & 63;
    if (i >= n) return;
    const long sbase = (long)blockIdx.y * 3 * n;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    const int e1 = rp[i + 1];
    int e = rp[i];
    for (; e + 1 < e1; e += 2) {  // two neighbours per step: two load chains in flight
        const int j0 = colx[e], j1 = colx[e + 1];
        const float* b0 = blk + 9L * e;
        const float* b1 = b0 + 9;
        const float* v0 = V + (sbase + 3L * j0) * C + ch;
        const float* v1 = V + (sbase + 3L * j1) * C + ch;
        const float x0 = v0[0], x1 = v0[C], x2 = v0[2 * C];
        const float y0 = v1[0], y1 = v1[C], y2 = v1[2 * C];
        a0 += b0[0] * x0 + b0[1] * x1 + b0[2] * x2 + b1[0] * y0 + b1[1] * y1 + b1[2] * y2;
        a1 += b0[3] * x0 + b0[4] * x1 + b0[5] * x2 + b1[3] * y0 + b1[4] * y1 + b1[5] * y2;
        a2 += b0[6] * x0 + b0[7] * x1 + b0[8] * x2 + b1[6] * y0 + b1[7] * y1 + b1[8] * y2;
    }
    if (e < e1) {
        const int j = colx[e];
        const float* b = blk + 9L * e;
        const float* vj = V + (sbase + 3L * j) * C + ch;
        const float x0 = vj[0], x1 = vj[C], x2 = vj[2 * C];
        a0 += b[0] * x0 + b[1] * x1 + b[2] * x2;
        a1 += b[3] * x0 + b[4] * x1 + b[5] * x2;
        a2 += b[6] * x0 + b[7] * x1 + b[8] * x2;
    }
    float* y = Y + (sbase + 3L * i) * C + ch;
    y[0] = a0;
    y[C] = a1;
    y[2 * C] = a2;
}

// Y = Ahat X on the 3 input features: one wave per node, lanes over the
// neighbours, fixed-order wave reduction.
__global__ __launch_bounds__(256) void k_spmm3(int n, const int* __restrict__ rp, const int* __restrict__ colx,
                                               const float* __restrict__ blk, const float* __restrict__ X,
                                               float* __restrict__ Y, const int* __restrict__ done)
{
    if (done != nullptr && *done) return;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const long sbase = (long)blockIdx.y * 3 * n;
    float a[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int e = rp[i] + lane; e < rp[i + 1]; e += 64) {
        const int j = colx[e];
        const float* b = blk + 9L * e;
        const float* xj = X + (sbase + 3L * j) * 3;
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int f = 0; f < 3; ++f)
#pragma unroll
                for (int g = 0; g < 3; ++g) a[3 * d + g] += b[3 * d + f] * xj[3 * f + g];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k] = wave_sum(a[k]);
    if (lane == 0) {
        float* y = Y + (sbase + 3L * i) * 3;
#pragma unroll
        for (int k = 0; k < 9; ++k) y[k] = a[k];
    }
}

// (Ahat V) of the layer's input comes precomputed in AV (k_spmm64, or
// k_spmm3 on X for the first layer): the kernel itself reads rows only.
template <bool FIRST, bool LAST, bool TRAIN>
__global__ __launch_bounds__(256) void k_layer(int n, const float* __restrict__ AV, const float* __restrict__ Vin,
                                               const float* __restrict__ X, const float* __restrict__ W_in,
                                               const uint16_t* __restrict__ Wt, float* __restrict__ Vout,
                                               const float* __restrict__ w_out, float* __restrict__ chat,
                                               const float* __restrict__ s_node, float* __restrict__ scaled,
                                               uint16_t* __restrict__ Uout, float* __restrict__ Yout, int linear,
                                               const int* __restrict__ done)
{
    __shared__ __align__(16) uint16_t sU[kRows * kUld];
    if (done != nullptr && *done) return;
    __shared__ __align__(16) uint16_t sW[C * kWld];
    __shared__ float sV[kRows * C];
    __shared__ float sPart[4 * kRows];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wv = t >> 6;
    const int node0 = blockIdx.x * kMpnnTile;
    const long sbase = (long)blockIdx.y * 3 * n;  // first row of this sample
    for (int q = t; q < C * 16; q += 256) {
        const int rr = q >> 4, cc = (q & 15) * 8;
        *reinterpret_cast<u32x4*>(&sW[rr * kWld + cc]) = *reinterpret_cast<const u32x4*>(&Wt[rr * 2 * C + cc]);
    }
    const int ch = lane;
    float wi0 = 0.f, wi1 = 0.f, wi2 = 0.f;
    if (FIRST) {
        wi0 = W_in[ch];
        wi1 = W_in[C + ch];
        wi2 = W_in[2 * C + ch];
    }
    for (int m = 0; m < kMpnnTile / 4; ++m) {
        const int li = wv + 4 * m;
        const int i = node0 + li;
        float v[3] = {0.f, 0.f, 0.f}, a[3] = {0.f, 0.f, 0.f};
        if (i < n) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const long row = sbase + 3L * i + d;
                if (FIRST) {
                    const float* xi = X + row * 3;
                    const float* ai = AV + row * 3;
                    v[d] = xi[0] * wi0 + xi[1] * wi1 + xi[2] * wi2;
                    a[d] = ai[0] * wi0 + ai[1] * wi1 + ai[2] * wi2;
                } else {
                    v[d] = Vin[row * C + ch];
                    a[d] = AV[row * C + ch];
                }
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int row = 3 * li + d;
            sU[row * kUld + ch] = f2bf(v[d]);
            sU[row * kUld + C + ch] = f2bf(a[d]);
            sV[row * C + ch] = v[d];
        }
    }
    __syncthreads();
    const int n0 = 16 * wv;
    const int lr = lane & 15;
    const int lk = (lane >> 4) * 8;
    const int col = n0 + lr;
    const float wo = LAST ? w_out[col] : 0.f;
#pragma unroll
    for (int mt = 0; mt < kRows / 16; ++mt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < (2 * C) / 32; ++kk) {
            const bf16x8 av = ld8(&sU[(16 * mt + lr) * kUld + 32 * kk + lk]);
            const bf16x8 bv = ld8(&sW[(n0 + lr) * kWld + 32 * kk + lk]);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = 16 * mt + (lane >> 4) * 4 + j;
            const int i = node0 + row / 3;
            const float y = acc[j];
            const float vo = sV[row * C + col] + (linear ? y : fmaxf(y, 0.f));
            if (i < n) {
                const long g = sbase + 3L * node0 + row;
                if (Vout) Vout[g * C + col] = vo;
                if (TRAIN) Yout[g * C + col] = y;
            }
            if (LAST) {
                float p = vo * wo;
                p += __shfl_xor(p, 1, 64);
                p += __shfl_xor(p, 2, 64);
                p += __shfl_xor(p, 4, 64);
                p += __shfl_xor(p, 8, 64);
                if (lr == 0) sPart[wv * kRows + row] = p;
            }
        }
    }
    if (TRAIN) {
        const int rows_valid = 3 * (n - node0 < kMpnnTile ? n - node0 : kMpnnTile);
        for (int q = t; q < rows_valid * 16; q += 256) {
            const int row = q >> 4, cc = (q & 15) * 8;
            *reinterpret_cast<u32x4*>(&Uout[(sbase + 3L * node0 + row) * 2 * C + cc]) =
                *reinterpret_cast<const u32x4*>(&sU[row * kUld + cc]);
        }
    }
    if (LAST) {
        __syncthreads();
        if (t < kRows) {
            const int i = node0 + t / 3;
            if (i < n) {
                const float c = ((sPart[t] + sPart[kRows + t]) + sPart[2 * kRows + t]) + sPart[3 * kRows + t];
                chat[sbase + 3L * node0 + t] = c;
                if (scaled) scaled[3L * node0 + t] = s_node[i] * c;
            }
        }
    }
}

// ---------------------------------------------------------------- backward
// dchat = (Ahat chat - q) / rows_total; loss partial sum(0.5 chat Ahat chat - chat q).
__global__ __launch_bounds__(kVec) void k_dchat(int n, int S, const int* __restrict__ rp,
                                                const int* __restrict__ colx, const float* __restrict__ blk,
                                                const float* __restrict__ chat, const float* __restrict__ q,
                                                float inv_rows, float* __restrict__ dchat, double* __restrict__ lpart)
{
    __shared__ double sm[16];
    const long u = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (u < (long)n * S) {
        const long smp = u / n;
        const int i = (int)(u - smp * n);
        const long sb = smp * 3L * n;
        float a[3] = {0.f, 0.f, 0.f};
        for (int e = rp[i]; e < rp[i + 1]; ++e) {
            const int j = colx[e];
            const float* b = blk + 9L * e;
            const float* cj = chat + sb + 3L * j;
            const float c0 = cj[0], c1 = cj[1], c2 = cj[2];
            a[0] += b[0] * c0 + b[1] * c1 + b[2] * c2;
            a[1] += b[3] * c0 + b[4] * c1 + b[5] * c2;
            a[2] += b[6] * c0 + b[7] * c1 + b[8] * c2;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const long row = sb + 3L * i + d;
            const float ci = chat[row], qi = q[row];
            dchat[row] = (a[d] - qi) * inv_rows;
            acc += 0.5 * (double)ci * (double)a[d] - (double)ci * (double)qi;
        }
    }
    const double s = block_sum(acc, sm);
    if (threadIdx.x == 0) lpart[blockIdx.x] = s;
}

// G = dchat (x) w_out; dw_out partials over 256-row chunks (one thread per channel).
__global__ __launch_bounds__(kVec) void k_dec_bwd(long rows, const float* __restrict__ dchat,
                                                  const float* __restrict__ VL, const float* __restrict__ w_out,
                                                  float* __restrict__ G, float* __restrict__ part)
{
    const int t = threadIdx.x;
    const int ch = t & 63;
    const int grp = t >> 6;
    const long r0 = (long)blockIdx.x * 256;
    float acc = 0.f;
    const float wo = w_out[ch];
    for (int k = grp; k < 256; k += 4) {
        const long row = r0 + k;
        if (row >= rows) break;
        const float dc = dchat[row];
        G[row * C + ch] = dc * wo;
        acc += VL[row * C + ch] * dc;
    }
    __shared__ float sm[4 * C];
    sm[grp * C + ch] = acc;
    __syncthreads();
    if (t < C) part[(long)blockIdx.x * C + t] = ((sm[t] + sm[C + t]) + sm[2 * C + t]) + sm[3 * C + t];
}

// dY = G * (Y > 0) (bf16, stored) and dU = dY W^T (float, 96 x 128 per workgroup).
__global__ __launch_bounds__(256) void k_bwd_dydu(int n, const float* __restrict__ G, const float* __restrict__ Y,
                                                  const uint16_t* __restrict__ Wn, uint16_t* __restrict__ dYout,
                                                  float* __restrict__ dU, int linear)
{
    __shared__ __align__(16) uint16_t sD[kRows * kDld];
    __shared__ __align__(16) uint16_t sW[2 * C * kDld];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wv = t >> 6;
    const int node0 = blockIdx.x * kMpnnTile;
    const long sbase = (long)blockIdx.y * 3 * n;
    const int rows_valid = 3 * (n - node0 < kMpnnTile ? n - node0 : kMpnnTile);
    // W (2C x C) as bf16 rows: B[k][col] = W[col][k] for dU = dY W^T
    for (int q = t; q < 2 * C * 8; q += 256) {
        const int rr = q >> 3, cc = (q & 7) * 8;
        *reinterpret_cast<u32x4*>(&sW[rr * kDld + cc]) = *reinterpret_cast<const u32x4*>(&Wn[rr * C + cc]);
    }
    for (int q = t; q < kRows * C; q += 256) {
        const int row = q / C, c = q % C;
        uint16_t h = 0;
        if (row < rows_valid) {
            const long g = (sbase + 3L * node0 + row) * C + c;
            const float d = (linear || Y[g] > 0.f) ? G[g] : 0.f;
            h = f2bf(d);
            dYout[g] = h;
        }
        sD[row * kDld + c] = h;
    }
    __syncthreads();
    const int lr = lane & 15;
    const int lk = (lane >> 4) * 8;
    // 6 row tiles x 8 column tiles; wave wv takes column tiles 2wv, 2wv+1
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
        const int n0 = 16 * (2 * wv + ct);
#pragma unroll
        for (int mt = 0; mt < kRows / 16; ++mt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kk = 0; kk < C / 32; ++kk) {
                const bf16x8 av = ld8(&sD[(16 * mt + lr) * kDld + 32 * kk + lk]);
                const bf16x8 bv = ld8(&sW[(n0 + lr) * kDld + 32 * kk + lk]);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = 16 * mt + (lane >> 4) * 4 + j;
                if (row < rows_valid) dU[(sbase + 3L * node0 + row) * 2 * C + n0 + lr] = acc[j];
            }
        }
    }
}

// G_new = G + dU[:, :C] + Ahat dU[:, C:] (one wave per node, lanes = channels).
__global__ __launch_bounds__(256) void k_bwd_gather(int n, const int* __restrict__ rp, const int* __restrict__ colx,
                                                    const float* __restrict__ blk, const float* __restrict__ G,
                                                    const float* __restrict__ dU, float* __restrict__ Gn)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int ch = threadIdx.x & 63;
    if (i >= n) return;
    const long sbase = (long)blockIdx.y * 3 * n;
    float a[3] = {0.f, 0.f, 0.f};
    for (int e = rp[i]; e < rp[i + 1]; ++e) {
        const int j = colx[e];
        const float* b = blk + 9L * e;
        const float* uj = dU + (sbase + 3L * j) * 2 * C + C + ch;
        const float x0 = uj[0], x1 = uj[2 * C], x2 = uj[4 * C];
        a[0] += b[0] * x0 + b[1] * x1 + b[2] * x2;
        a[1] += b[3] * x0 + b[4] * x1 + b[5] * x2;
        a[2] += b[6] * x0 + b[7] * x1 + b[8] * x2;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const long row = sbase + 3L * i + d;
        Gn[row * C + ch] = G[row * C + ch] + dU[row * 2 * C + ch] + a[d];
    }
}

// First layer: dW_in[g][c] partial = sum_rows X[row][g] (G + dU_self)[row][c]
//                                    + (Ahat X)[row][g] dU_nb[row][c]
__global__ __launch_bounds__(256) void k_bwd_win(int n, const int* __restrict__ rp, const int* __restrict__ colx,
                                                 const float* __restrict__ blk, const float* __restrict__ X,
                                                 const float* __restrict__ G, const float* __restrict__ dU,
                                                 float* __restrict__ part)
{
    __shared__ float sm[4 * 3 * C];
    const int t = threadIdx.x;
    const int ch = t & 63;
    const int wv = t >> 6;
    const int node0 = blockIdx.x * kMpnnTile;
    const long sbase = (long)blockIdx.y * 3 * n;
    float acc[3] = {0.f, 0.f, 0.f};
    for (int m = 0; m < kMpnnTile / 4; ++m) {
        const int i = node0 + wv + 4 * m;
        if (i >= n) break;
        float ax[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int e = rp[i]; e < rp[i + 1]; ++e) {
            const int j = colx[e];
            const float* b = blk + 9L * e;
            const float* xj = X + (sbase + 3L * j) * 3;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int f = 0; f < 3; ++f)
#pragma unroll
                    for (int g = 0; g < 3; ++g) ax[3 * d + g] += b[3 * d + f] * xj[3 * f + g];
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const long row = sbase + 3L * i + d;
            const float gs = G[row * C + ch] + dU[row * 2 * C + ch];
            const float gn = dU[row * 2 * C + C + ch];
#pragma unroll
            for (int g = 0; g < 3; ++g) acc[g] += X[row * 3 + g] * gs + ax[3 * d + g] * gn;
        }
    }
#pragma unroll
    for (int g = 0; g < 3; ++g) sm[(wv * 3 + g) * C + ch] = acc[g];
    __syncthreads();
    if (t < 3 * C) {
        const int g = t / C, c = t % C;
        const long b = (long)blockIdx.y * gridDim.x + blockIdx.x;
        part[b * 3 * C + t] = ((sm[(0 * 3 + g) * C + c] + sm[(1 * 3 + g) * C + c]) + sm[(2 * 3 + g) * C + c]) +
                              sm[(3 * 3 + g) * C + c];
    }
}

// dW partial (2C x C) = U^T dY over kDwRows rows per workgroup.
__global__ __launch_bounds__(256) void k_dw(long rows, const uint16_t* __restrict__ U, const uint16_t* __restrict__ dY,
                                            float* __restrict__ part)
{
    __shared__ __align__(16) uint16_t sUT[2 * C * (32 + 8)];  // transposed chunk: [u][row]
    __shared__ __align__(16) uint16_t sDT[C * (32 + 8)];      // [c][row]
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wv = t >> 6;
    const int lr = lane & 15;
    const int lk = (lane >> 4) * 8;
    const long r0 = (long)blockIdx.x * kDwRows;
    f32x4 acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < kDwRows; k0 += 32) {
        __syncthreads();
        for (int q = t; q < 32 * 2 * C; q += 256) {
            const int row = q / (2 * C), u = q % (2 * C);
            const long g = r0 + k0 + row;
            sUT[u * 40 + row] = g < rows ? U[g * 2 * C + u] : (uint16_t)0;
        }
        for (int q = t; q < 32 * C; q += 256) {
            const int row = q / C, c = q % C;
            const long g = r0 + k0 + row;
            sDT[c * 40 + row] = g < rows ? dY[g * C + c] : (uint16_t)0;
        }
        __syncthreads();
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int m0 = 16 * (2 * wv + a);
            const bf16x8 av = ld8(&sUT[(m0 + lr) * 40 + lk]);
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const bf16x8 bv = ld8(&sDT[(16 * b + lr) * 40 + lk]);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[a][b], 0, 0, 0);
            }
        }
    }
    float* out = part + (long)blockIdx.x * 2 * C * C;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int u = 16 * (2 * wv + a) + (lane >> 4) * 4 + j;
                out[u * C + 16 * b + lr] = acc[a][b][j];
            }
}

// out[p] (+)= sum_b part[b][p], b in fixed order.
__global__ void k_reduce_parts(const float* __restrict__ part, int nb, int np, float* __restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += (double)part[(long)b * np + p];
    out[p] = (float)s;
}

__global__ void k_adam(int np, float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const float gi = g[p];
    const float mi = b1 * m[p] + (1.f - b1) * gi;
    const float vi = b2 * v[p] + (1.f - b2) * gi * gi;
    m[p] = mi;
    v[p] = vi;
    w[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// per-layer bf16 copies: Wt[l] = W_l^T (C x 2C), Wn[l] = W_l (2C x C)
__global__ void k_to_bf16(int L, const float* __restrict__ w, uint16_t* __restrict__ wt, uint16_t* __restrict__ wn)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= L * 2 * C * C) return;
    const int l = p / (2 * C * C), q = p % (2 * C * C);
    const int u = q / C, c = q % C;
    const uint16_t h = f2bf(w[3 * C + p]);
    wn[p] = h;
    wt[(long)l * 2 * C * C + c * 2 * C + u] = h;
}

int g256(long n) { return (int)((n + 255) / 256); }

}  // namespace

// ----------------------------------------------------------------- graph
void CoarseGraph::build(hipStream_t st, const double* vals, const int* rows, const int* cols, int nb, int n_nodes)
{
    SGIPC_CHECK(n_nodes >= 1 && nb >= 1, SGIPC_EMPTY, "empty level-1 operator");
    n = n_nodes;
    const long slots = 2L * nb;
    keys_.reserve(slots);
    skeys_.reserve(slots);
    idx_.reserve(slots);
    sidx_.reserve(slots);
    k_g_keys<<<g256(nb), 256, 0, st>>>(rows, cols, nb, keys_.get(), idx_.get());
    SGIPC_LAUNCH_CHECK();
    int bits = 1;
    while ((1L << bits) < n) ++bits;
    size_t need = 0;
    // the unused slots carry ~0 and must stay last: sort all 64 bits
    SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, need, keys_.get(), skeys_.get(), idx_.get(), sidx_.get(),
                                        (size_t)slots, 0, 64, st));
    temp_.reserve(need);
    SGIPC_HIP(rocprim::radix_sort_pairs(temp_.get(), need, keys_.get(), skeys_.get(), idx_.get(), sidx_.get(),
                                        (size_t)slots, 0, 64, st));
    // E = 2 nb - (diagonal blocks)
    cnt_.reserve(1);
    SGIPC_HIP(hipMemsetAsync(cnt_.get(), 0, sizeof(int), st));
    k_g_count_diag<<<g256(nb), 256, 0, st>>>(rows, cols, nb, cnt_.get());
    SGIPC_LAUNCH_CHECK();
    int nd = 0;
    SGIPC_HIP(hipMemcpyAsync(&nd, cnt_.get(), sizeof(int), hipMemcpyDeviceToHost, st));
    SGIPC_HIP(hipStreamSynchronize(st));
    nnz = (int)(slots - nd);
    row_ptr.reserve(n + 1);
    col.reserve(nnz);
    blk.reserve(9L * nnz);
    s.reserve(n);
    SGIPC_HIP(hipMemsetAsync(s.get(), 0, sizeof(float) * n, st));
    k_g_rowptr<<<g256(nnz + 1), 256, 0, st>>>(skeys_.get(), nnz, n, row_ptr.get());
    SGIPC_LAUNCH_CHECK();
    k_g_scale<<<g256(nnz), 256, 0, st>>>(skeys_.get(), sidx_.get(), nnz, vals, s.get());
    SGIPC_LAUNCH_CHECK();
    k_g_blocks<<<g256(nnz), 256, 0, st>>>(skeys_.get(), sidx_.get(), nnz, vals, s.get(), col.get(), blk.get());
    SGIPC_LAUNCH_CHECK();
}

void CoarseGraph::copy_from(hipStream_t st, const CoarseGraph& o)
{
    n = o.n;
    nnz = o.nnz;
    row_ptr.reserve(n + 1);
    col.reserve(nnz);
    blk.reserve(9L * nnz);
    s.reserve(n);
    SGIPC_HIP(hipMemcpyAsync(row_ptr.get(), o.row_ptr.get(), sizeof(int) * (n + 1), hipMemcpyDeviceToDevice, st));
    SGIPC_HIP(hipMemcpyAsync(col.get(), o.col.get(), sizeof(int) * nnz, hipMemcpyDeviceToDevice, st));
    SGIPC_HIP(hipMemcpyAsync(blk.get(), o.blk.get(), sizeof(float) * 9L * nnz, hipMemcpyDeviceToDevice, st));
    SGIPC_HIP(hipMemcpyAsync(s.get(), o.s.get(), sizeof(float) * n, hipMemcpyDeviceToDevice, st));
}

// ----------------------------------------------------------------- network
Mpnn::Mpnn(Context* c, int layers, int linear) : ctx(c), L(layers), linear_(linear ? 1 : 0)
{
    SGIPC_CHECK(layers >= 1 && layers <= kMpnnMaxLayers, SGIPC_BAD_ARG, "layer count out of range");
    const int np = n_params();
    w_.reserve(np);
    m_.reserve(np);
    v_.reserve(np);
    grad_.reserve(np);
    wt_.reserve((size_t)L * 2 * C * C);
    wn_.reserve((size_t)L * 2 * C * C);
    init_random(0);
}

void Mpnn::init_random(uint64_t seed)
{
    // He-style normal initial weights; the output layer starts small so
    // the first corrections are close to zero
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<float> h(n_params());
    size_t p = 0;
    for (int k = 0; k < 3 * C; ++k) h[p++] = (float)(nd(rng) * std::sqrt(1.0 / 3.0));
    for (int k = 0; k < L * 2 * C * C; ++k) h[p++] = (float)(nd(rng) * std::sqrt(1.0 / (2.0 * C)) * 0.5);
    for (int k = 0; k < C; ++k) h[p++] = (float)(nd(rng) * std::sqrt(1.0 / C) * 0.1);
    set_params(h.data());
}

void Mpnn::set_params(const float* host)
{
    SGIPC_HIP(hipMemcpyAsync(w_.get(), host, sizeof(float) * n_params(), hipMemcpyHostToDevice, ctx->stream));
    SGIPC_HIP(hipStreamSynchronize(ctx->stream));
    dirty_ = true;
    reset_optimizer();
}

void Mpnn::get_params(float* host) const
{
    SGIPC_HIP(hipMemcpyAsync(host, w_.get(), sizeof(float) * n_params(), hipMemcpyDeviceToHost, ctx->stream));
    SGIPC_HIP(hipStreamSynchronize(ctx->stream));
}

void Mpnn::reset_optimizer()
{
    launch_fill<float>(ctx->stream, m_.get(), 0.f, n_params());
    launch_fill<float>(ctx->stream, v_.get(), 0.f, n_params());
    step_ = 0;
}

void Mpnn::sync_bf16(hipStream_t st)
{
    if (!dirty_) return;
    k_to_bf16<<<g256((long)L * 2 * C * C), 256, 0, st>>>(L, w_.get(), wt_.get(), wn_.get());
    SGIPC_LAUNCH_CHECK();
    dirty_ = false;
}

void Mpnn::reserve_train(long rows, int tiles_total)
{
    (void)tiles_total;
    if (rows <= rows_cap_) return;
    for (int l = 0; l <= L; ++l) V_[l].reserve((size_t)rows * C);
    for (int l = 0; l < L; ++l) {
        U_[l].reserve((size_t)rows * 2 * C);
        Y_[l].reserve((size_t)rows * C);
    }
    G_.reserve((size_t)rows * C);
    G2_.reserve((size_t)rows * C);
    dU_.reserve((size_t)rows * 2 * C);
    dY_.reserve((size_t)rows * C);
    chat_.reserve(rows);
    dchat_.reserve(rows);
    rows_cap_ = rows;
}

void Mpnn::forward(hipStream_t st, const CoarseGraph& g, const float* X, int S, float* chat, float* scaled,
                   const int* done)
{
    SGIPC_CHECK(g.n >= 1 && S >= 1, SGIPC_EMPTY, "empty graph or batch");
    SGIPC_CHECK(scaled == nullptr || S == 1, SGIPC_BAD_ARG, "scaled output needs a single sample");
    sync_bf16(st);
    const long rows = 3L * g.n * S;
    V_[0].reserve((size_t)rows * C);
    V_[1].reserve((size_t)rows * C);
    AV_.reserve((size_t)rows * C);
    const dim3 grid((g.n + kMpnnTile - 1) / kMpnnTile, S);
    const dim3 gw((g.n + 3) / 4, S);  // one wave per node
    const float* w_in = w_.get();
    const float* w_out = w_.get() + 3 * C + (long)L * 2 * C * C;
    const float* vin = nullptr;
    for (int l = 0; l < L; ++l) {
        const uint16_t* wt = wt_.get() + (long)l * 2 * C * C;
        float* vout = V_[(l + 1) & 1].get();
        const bool first = l == 0, last = l == L - 1;
        if (first) k_spmm3<<<gw, 256, 0, st>>>(g.n, g.row_ptr.get(), g.col.get(), g.blk.get(), X, AV_.get(), done);
        else k_spmm64<<<gw, 256, 0, st>>>(g.n, g.row_ptr.get(), g.col.get(), g.blk.get(), vin, AV_.get());
        SGIPC_LAUNCH_CHECK();
#define SGIPC_LAYER(F, LA)                                                                                     \
    k_layer<F, LA, false><<<grid, 256, 0, st>>>(g.n, AV_.get(), vin, X, w_in, wt, LA ? nullptr : vout, w_out, chat, \
                                                g.s.get(), scaled, nullptr, nullptr, linear_, done)
        if (first && last) SGIPC_LAYER(true, true);
        else if (first) SGIPC_LAYER(true, false);
        else if (last) SGIPC_LAYER(false, true);
        else SGIPC_LAYER(false, false);
#undef SGIPC_LAYER
        SGIPC_LAUNCH_CHECK();
        vin = vout;
    }
}

double Mpnn::loss_grad(hipStream_t st, const CoarseGraph& g, const float* X, const float* q, int S, float* grad)
{
    SGIPC_CHECK(g.n >= 1 && S >= 1, SGIPC_EMPTY, "empty graph or batch");
    sync_bf16(st);
    const int n = g.n;
    const long rows = 3L * n * S;
    const int tiles = (n + kMpnnTile - 1) / kMpnnTile;
    reserve_train(rows, tiles * S);
    const dim3 grid(tiles, S);
    const dim3 gw((n + 3) / 4, S);  // one wave per node
    const float* w_in = w_.get();
    const float* w_out = w_.get() + 3 * C + (long)L * 2 * C * C;
    AV_.reserve((size_t)rows * C);
    // forward, keeping U, Y and every V
    for (int l = 0; l < L; ++l) {
        const uint16_t* wt = wt_.get() + (long)l * 2 * C * C;
        const bool first = l == 0, last = l == L - 1;
        if (first) k_spmm3<<<gw, 256, 0, st>>>(n, g.row_ptr.get(), g.col.get(), g.blk.get(), X, AV_.get(), nullptr);
        else k_spmm64<<<gw, 256, 0, st>>>(n, g.row_ptr.get(), g.col.get(), g.blk.get(), V_[l].get(), AV_.get());
        SGIPC_LAUNCH_CHECK();
#define SGIPC_LAYER(F, LA)                                                                                   \
    k_layer<F, LA, true><<<grid, 256, 0, st>>>(n, AV_.get(), V_[l].get(), X, w_in, wt, V_[l + 1].get(), w_out, \
                                               chat_.get(), g.s.get(), nullptr, U_[l].get(), Y_[l].get(),   \
                                               linear_, nullptr)
        if (first && last) SGIPC_LAYER(true, true);
        else if (first) SGIPC_LAYER(true, false);
        else if (last) SGIPC_LAYER(false, true);
        else SGIPC_LAYER(false, false);
#undef SGIPC_LAYER
        SGIPC_LAUNCH_CHECK();
    }
    // loss and dchat
    const long nodes = (long)n * S;
    const int gl = g256(nodes);
    const int n_dec = g256(rows);
    const int n_dw = (int)((rows + kDwRows - 1) / kDwRows);
    const int n_win = tiles * S;
    const long part_need = std::max({(long)gl, (long)n_dec * C, (long)n_dw * 2 * C * C, (long)n_win * 3 * C});
    part_.reserve(part_need);
    DevBuf<double>& lp = lpart_;
    lp.reserve(gl);
    k_dchat<<<gl, kVec, 0, st>>>(n, S, g.row_ptr.get(), g.col.get(), g.blk.get(), chat_.get(), q,
                                 1.0f / (float)rows, dchat_.get(), lp.get());
    SGIPC_LAUNCH_CHECK();
    std::vector<double> h_lp(gl);
    SGIPC_HIP(hipMemcpyAsync(h_lp.data(), lp.get(), sizeof(double) * gl, hipMemcpyDeviceToHost, st));
    // decoder
    k_dec_bwd<<<n_dec, kVec, 0, st>>>(rows, dchat_.get(), V_[L].get(), w_out, G_.get(), part_.get());
    SGIPC_LAUNCH_CHECK();
    float* g_win = grad;
    float* g_wout = grad + 3 * C + (long)L * 2 * C * C;
    k_reduce_parts<<<1, C, 0, st>>>(part_.get(), n_dec, C, g_wout);
    SGIPC_LAUNCH_CHECK();
    float* Gc = G_.get();
    float* Gn = G2_.get();
    for (int l = L - 1; l >= 0; --l) {
        const uint16_t* wn = wn_.get() + (long)l * 2 * C * C;
        k_bwd_dydu<<<grid, 256, 0, st>>>(n, Gc, Y_[l].get(), wn, dY_.get(), dU_.get(), linear_);
        SGIPC_LAUNCH_CHECK();
        k_dw<<<n_dw, 256, 0, st>>>(rows, U_[l].get(), dY_.get(), part_.get());
        SGIPC_LAUNCH_CHECK();
        k_reduce_parts<<<g256(2 * C * C), 256, 0, st>>>(part_.get(), n_dw, 2 * C * C,
                                                         grad + 3 * C + (long)l * 2 * C * C);
        SGIPC_LAUNCH_CHECK();
        if (l > 0) {
            k_bwd_gather<<<gw, 256, 0, st>>>(n, g.row_ptr.get(), g.col.get(), g.blk.get(), Gc, dU_.get(), Gn);
            SGIPC_LAUNCH_CHECK();
            std::swap(Gc, Gn);
        } else {
            k_bwd_win<<<grid, 256, 0, st>>>(n, g.row_ptr.get(), g.col.get(), g.blk.get(), X, Gc, dU_.get(),
                                            part_.get());
            SGIPC_LAUNCH_CHECK();
            k_reduce_parts<<<1, 3 * C, 0, st>>>(part_.get(), n_win, 3 * C, g_win);
            SGIPC_LAUNCH_CHECK();
        }
    }
    SGIPC_HIP(hipStreamSynchronize(st));
    double loss = 0.0;
    for (int b = 0; b < gl; ++b) loss += h_lp[b];
    return loss / (double)rows;
}

void Mpnn::adam_step(hipStream_t st, float lr, float b1, float b2, float eps)
{
    ++step_;
    const float c1 = 1.f - std::pow(b1, (float)step_);
    const float c2 = 1.f - std::pow(b2, (float)step_);
    k_adam<<<g256(n_params()), 256, 0, st>>>(n_params(), w_.get(), m_.get(), v_.get(), grad_.get(), lr, b1, b2, eps,
                                             c1, c2);
    SGIPC_LAUNCH_CHECK();
    dirty_ = true;
}

}  // namespace sgipc
