// Tile-symmetric SELL SpMV for the PCG loop on gfx950.
//
// The reference streams the upper BCOO and scatters the transposed half with
// atomics (Spmv::srbk_spmv, StiffGIPC/linear_system/utils/spmv.cu:188-313).
// Here the matrix is re-laid out once per Newton step so that each upper
// block is read from HBM once per SpMV without atomics:
//
//   * rows are cut into tiles of kTile consecutive (Morton-ordered) rows; one
//     workgroup owns a tile and writes every row of it exactly once;
//   * an upper block (i, j) with both ends in the same tile is stored once, in
//     row i.  Its H x_j goes to row i in registers, its H^T x_i to a slot of
//     the tile's LDS array, numbered in (j, i) order, that row j adds up after
//     a barrier (fixed order, so the result is deterministic);
//   * a block whose ends lie in different tiles is stored twice, as H in row i
//     and as H^T in row j (about a fifth of the blocks on Morton-ordered
//     meshes with 512-row tiles);
//   * inside a tile the rows are sorted by entry count and packed as SELL-64:
//     slot s of lane l holds value c at ((S_q + s) * 9 + c) * 64 + l, so each
//     of the 9 value loads of a slot is one contiguous 512-byte access; a
//     lane past its row's end reads lane 0's address of the same slot (already
//     fetched) and masks the product, so padding costs no HBM bytes;
//   * two waves share a slice, each taking half of its slots, and combine the
//     two partial rows through LDS.
// Per tile: kTile rows, kSplit * kTile threads; x^T(Ax) partials per tile.
//
// Halo form (used whenever its LDS fits): after packing, every column word
// that is not an in-tile transposed slot is rewritten as an index into the
// tile's local copy of the gathered vector -- the tile's own rows, then its
// halo, the distinct outside vertices its stored blocks reference (ascending).
// The local copy is staged in LDS before the slots are streamed, so the slot
// loop reads no vector entry from memory: the global gather of a slot's
// column sat between that slot's loads and its products and cost one more
// memory round trip per slot.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include <hip/hip_ext.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "spmv_full.hpp"
#include "wave_ops.hpp"

namespace sgipc {

namespace {

constexpr int kTile = 512;                 // rows per workgroup
constexpr int kSplit = 2;                  // waves per slice
constexpr int kThreads = kTile * kSplit;   // 1024
constexpr int kTcap = 4096;                // in-tile transposed slots per tile (LDS)
constexpr uint32_t kInTile = 0x80000000u;
constexpr int kWaves = kThreads / 64;          // waves of a workgroup
constexpr int kSlices = kTile / 64;            // SELL-64 slices of a tile
constexpr int kPieces = kWaves + kSlices - 1;  // (wave, slice) partial rows of the halo kernel
constexpr int kLdsBytes = 163840;              // LDS of one CU on gfx950
constexpr int kLongRow = 128;                  // rows with more blocks are kept outside the tiles

// Workgroups are handed to the 8 XCDs of the part round-robin in blockIdx
// order.  Mapping block b to tile start(b % 8) + b / 8 gives every XCD a
// contiguous run of tiles, so the x entries neighbouring tiles share are
// fetched into one XCD's L2 instead of into all of them.  A bijection on
// [0, n): XCD x owns n / 8 + (x < n % 8) tiles starting at x (n / 8) + min(x, n % 8).
__device__ __forceinline__ int xcd_tile(int b, int n)
{
    const int x = b & 7, q = n >> 3, rem = n & 7;
    return x * q + (x < rem ? x : rem) + (b >> 3);
}

// With kFuseP the gathered vector is p_new = z + beta p_old, formed on the fly
// (beta = rz_new / rz_old from partials, summed in a fixed order), and every
// owned row of p_new is written: the PCG p update rides along with the SpMV.
template <typename TM, typename TV, bool kDot, bool kFuseP>
__global__ __launch_bounds__(kThreads) void k_spmv_tsym(const TM* __restrict__ vals, const uint32_t* __restrict__ cw,
                                                        const int* __restrict__ slice_ptr,
                                                        const int* __restrict__ slice_w, const int* __restrict__ perm,
                                                        const int* __restrict__ plen,
                                                        const uint32_t* __restrict__ tinfo, TV a,
                                                        const TV* __restrict__ x, TV b, TV* __restrict__ y,
                                                        double* __restrict__ dot_part, const int* __restrict__ done,
                                                        const TV* __restrict__ zf, TV* __restrict__ p_new,
                                                        const double* __restrict__ rzn,
                                                        const double* __restrict__ rzo, int n_rz, int n_rows)
{
    __shared__ TV tl[kTcap * 3];
    __shared__ TV red[3 * kTile];
    __shared__ TV xs[3 * kTile];  // the gathered vector over this tile's own rows
    __shared__ double dred[16];
    constexpr int kU = sizeof(TM) == 4 ? 3 : 2;  // slots per step (loads in flight)
    if (done != nullptr && *done) return;
    double beta = 0.0;
    if (kFuseP) beta = block_sum_array(rzn, n_rz, dred) / block_sum_array(rzo, n_rz, dred);
    // gathered value: x itself, or p_new = z + beta p_old
    auto gx = [&](long k) -> TV {
        if (kFuseP) return (TV)((double)zf[k] + beta * (double)x[k]);
        return x[k];
    };
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    {
        // in-tile columns read x from LDS: one coalesced load of the tile's rows
        const long b3 = 3L * tile * kTile, e3 = 3L * n_rows;
        for (int t = threadIdx.x; t < 3 * kTile; t += kThreads) xs[t] = b3 + t < e3 ? gx(b3 + t) : TV(0);
        __syncthreads();
    }
    const int part = threadIdx.x / kTile;
    const int r = threadIdx.x % kTile;
    const int lane = r & 63;
    const int q = tile * (kTile / 64) + (r >> 6);
    const int p = q * 64 + lane;
    const int i = perm[p];
    const int len = plen[p];
    TV xi0 = 0, xi1 = 0, xi2 = 0;
    if (i >= 0) {
        xi0 = gx(3 * (long)i);
        xi1 = gx(3 * (long)i + 1);
        xi2 = gx(3 * (long)i + 2);
        if (kFuseP && part == 0) {
            p_new[3 * (long)i] = xi0;
            p_new[3 * (long)i + 1] = xi1;
            p_new[3 * (long)i + 2] = xi2;
        }
    }
    const long S = slice_ptr[q];
    const int W = slice_w[q];
    const int s1 = (W * (part + 1)) / kSplit;
    int s = (W * part) / kSplit;
    TV a0 = 0, a1 = 0, a2 = 0;
    const uint32_t* cp = cw + S * 64;
    const TM* vp = vals + S * 9 * 64;
    // one slot: its block times x_j into the row sums, and its transposed
    // product into the LDS slot when the block is in-tile
    auto slot = [&](const TM* h, uint32_t w, bool act) {
        const bool in = act && (w & kInTile);
        const TV m = act ? TV(1) : TV(0);
        TV x0, x1, x2;
        if (w & kInTile) {
            const int lc = 3 * ((w >> 16) & 0x7FFF);
            x0 = m * xs[lc];
            x1 = m * xs[lc + 1];
            x2 = m * xs[lc + 2];
        } else {
            const long j = (long)w;
            x0 = m * gx(3 * j);
            x1 = m * gx(3 * j + 1);
            x2 = m * gx(3 * j + 2);
        }
        a0 += (TV)h[0] * x0 + (TV)h[3] * x1 + (TV)h[6] * x2;
        a1 += (TV)h[1] * x0 + (TV)h[4] * x1 + (TV)h[7] * x2;
        a2 += (TV)h[2] * x0 + (TV)h[5] * x1 + (TV)h[8] * x2;
        if (in) {
            const int ts = w & 0xFFFF;
            tl[3 * ts] = (TV)h[0] * xi0 + (TV)h[1] * xi1 + (TV)h[2] * xi2;
            tl[3 * ts + 1] = (TV)h[3] * xi0 + (TV)h[4] * xi1 + (TV)h[5] * xi2;
            tl[3 * ts + 2] = (TV)h[6] * xi0 + (TV)h[7] * xi1 + (TV)h[8] * xi2;
        }
    };
    // kU slots per step, all their loads issued first; the matrix stream is
    // read once, so its loads are non-temporal and leave the L2 to x
    for (; s + kU <= s1; s += kU) {
        TM h[kU][9];
        uint32_t w[kU];
        bool act[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            act[u] = s + u < len;
            const int l = act[u] ? lane : 0;
            w[u] = __builtin_nontemporal_load(cp + (long)(s + u) * 64 + l);
#pragma unroll
            for (int c = 0; c < 9; ++c) h[u][c] = __builtin_nontemporal_load(vp + ((long)(s + u) * 9 + c) * 64 + l);
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) slot(h[u], w[u], act[u]);
    }
    for (; s < s1; ++s) {
        const bool act0 = s < len;
        const int l0 = act0 ? lane : 0;
        TM h[9];
        const uint32_t w0 = __builtin_nontemporal_load(cp + (long)s * 64 + l0);
#pragma unroll
        for (int c = 0; c < 9; ++c) h[c] = __builtin_nontemporal_load(vp + ((long)s * 9 + c) * 64 + l0);
        slot(h, w0, act0);
    }
    if (part != 0) {
        red[r] = a0;
        red[kTile + r] = a1;
        red[2 * kTile + r] = a2;
    }
    __syncthreads();
    double dot = 0.0;
    if (part == 0 && i >= 0) {
        a0 += red[r];
        a1 += red[kTile + r];
        a2 += red[2 * kTile + r];
        const uint32_t ti = tinfo[i];
        const int tb = (int)(ti >> 16), tc = (int)(ti & 0xFFFF);
        for (int k = tb; k < tb + tc; ++k) {
            a0 += tl[3 * k];
            a1 += tl[3 * k + 1];
            a2 += tl[3 * k + 2];
        }
        TV* yi = y + 3 * (long)i;
        if (b == TV(0)) {
            yi[0] = a * a0;
            yi[1] = a * a1;
            yi[2] = a * a2;
        } else {
            yi[0] = a * a0 + b * yi[0];
            yi[1] = a * a1 + b * yi[1];
            yi[2] = a * a2 + b * yi[2];
        }
        if (kDot) dot = (double)xi0 * (double)a0 + (double)xi1 * (double)a1 + (double)xi2 * (double)a2;
    }
    if (kDot) {
        const double sblk = block_sum(dot, dred);
        if (threadIdx.x == 0) dot_part[tile] = sblk;
    }
}

// Halo kernel.  The tile's slots (its slices back to back) are cut into kWaves
// equal ranges, one per wave, across slice boundaries; a wave parks one partial
// row sum per slice it touches in LDS piece wave + slice (distinct for distinct
// non-empty pieces, as both cut lists increase) and each row adds its pieces in
// wave order.  Each step's loads are issued before the previous step is used.
// Tiles run in `order` (heaviest first inside each XCD's range), so the
// lightest tiles fill the second round of workgroups.
// Dynamic LDS: tl[3 tcap], piece[3 * 64 kPieces], xl[3 (kTile + hcap)] of TV.
// Launch shape and LDS element type of the halo kernel per (matrix, vector)
// type: the FP64 kernel keeps exact FP64 copies in LDS (one 1024-thread
// workgroup per CU); the kernels with FP32 values keep FP32 copies (the
// vector and the in-tile transposed products rounded to FP32, accumulation
// in the vector type) in 512-thread workgroups, two of which fit a CU's LDS
// and registers, so one workgroup's staging and reduction overlap the other's
// stream.
template <typename TM, typename TV>
struct HaloCfg {
    static constexpr int NT = 1024;
    using TL = TV;
};
template <>
struct HaloCfg<float, double> {
    static constexpr int NT = 512;
    using TL = float;
};
template <>
struct HaloCfg<float, float> {
    static constexpr int NT = 512;
    using TL = float;
};

template <typename TM, typename TV, bool kDot, bool kFuseP, bool kSum1 = false>
__global__ __launch_bounds__((HaloCfg<TM, TV>::NT), 4) void k_spmv_halo(const TM* __restrict__ vals, const uint32_t* __restrict__ cw,
                                                        const int* __restrict__ slice_ptr,
                                                        const int* __restrict__ perm, const int* __restrict__ plen,
                                                        const uint32_t* __restrict__ tinfo,
                                                        const int* __restrict__ halo_ptr,
                                                        const int* __restrict__ halo_list,
                                                        const int* __restrict__ order, int tcap, TV a,
                                                        const TV* __restrict__ x, TV b, TV* __restrict__ y,
                                                        double* __restrict__ dot_part, const int* __restrict__ done,
                                                        const TV* __restrict__ zf, TV* __restrict__ p_new,
                                                        const double* __restrict__ rzn,
                                                        const double* __restrict__ rzo, int n_rz, int n_rows,
                                                        Level1Sums l1)
{
    using TL = typename HaloCfg<TM, TV>::TL;
    constexpr int kThreads = HaloCfg<TM, TV>::NT;
    constexpr int kWaves = kThreads / 64;
    constexpr int kPieces = kWaves + kSlices - 1;
    constexpr bool kExactLds = std::is_same<TL, TV>::value;
    extern __shared__ __align__(16) unsigned char lds_dyn[];
    TL* const tl = reinterpret_cast<TL*>(lds_dyn);
    TL* const piece = tl + 3 * tcap;
    TL* const xl = piece + kPieces * 3 * 64;
    __shared__ int s_len[kTile];
    __shared__ int s_row[kTile];  // row of each tile position, relative to the tile (-1: padding)
    __shared__ int s_sb[kSlices + 1];
    __shared__ double dred[16];
    // slots per step (one for the fused p update, whose staging needs more registers)
    constexpr int kU = (sizeof(TM) == 4 && !kFuseP) ? 2 : 1;
    if (done != nullptr && *done) return;
    double beta = 0.0;
    if (kFuseP) beta = block_sum_array(rzn, n_rz, dred) / block_sum_array(rzo, n_rz, dred);
    auto gx = [&](long k) -> TV {
        if (kFuseP) return (TV)((double)zf[k] + beta * (double)x[k]);
        return x[k];
    };
    const int tile = order[xcd_tile(blockIdx.x, gridDim.x)];
    const long b3 = 3L * tile * kTile;
    const int q0 = tile * kSlices;
    const long A = slice_ptr[q0];
    const int tot = (int)((long)slice_ptr[q0 + kSlices] - A);
    uint32_t ti_r = 0;
    // row table and slice bounds first, so the first matrix loads can go out
    // before the vector is staged (its loads then overlap theirs)
    if (threadIdx.x < kTile) {
        const int pr = perm[q0 * 64 + threadIdx.x];
        s_row[threadIdx.x] = pr >= 0 ? pr - tile * kTile : -1;
        s_len[threadIdx.x] = plen[q0 * 64 + threadIdx.x];
        if (pr >= 0) ti_r = tinfo[pr];
    }
    if (threadIdx.x <= kSlices) s_sb[threadIdx.x] = (int)((long)slice_ptr[q0 + threadIdx.x] - A);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int w0 = (int)((long)tot * wave / kWaves), w1 = (int)((long)tot * (wave + 1) / kWaves);
    const uint32_t* cp = cw + A * 64;
    const TM* vp = vals + A * 9 * 64;
    // the loads of the step at local slot ps of slice sl (matrix stream read
    // once: non-temporal); slots past the wave's range repeat slot ps, masked
    auto issue = [&](int ps, int sl, uint32_t* W, TM (*H)[9], bool* act) {
        const int send = min(w1, s_sb[sl + 1]);
        const int len = s_len[sl * 64 + lane];
        const int s0 = ps - s_sb[sl];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const bool in = ps + u < send;
            act[u] = in && s0 + u < len;
            const int l = act[u] ? lane : 0;
            const long g = in ? ps + u : ps;
            W[u] = __builtin_nontemporal_load(cp + g * 64 + l);
#pragma unroll
            for (int c = 0; c < 9; ++c) H[u][c] = __builtin_nontemporal_load(vp + (g * 9 + c) * 64 + l);
        }
    };
    int sq = 0, pos = w0;
    uint32_t Wa[kU];
    TM Ha[kU][9];
    bool aa[kU];
    if (w0 < w1) {
        while (s_sb[sq + 1] <= w0) ++sq;
        issue(pos, sq, Wa, Ha, aa);
    }
    {
        // stage the tile's own rows and its halo
        const long e3 = 3L * n_rows;
        for (int t = threadIdx.x; t < 3 * kTile; t += kThreads) xl[t] = (TL)(b3 + t < e3 ? gx(b3 + t) : TV(0));
        const int h0 = halo_ptr[tile], hn = halo_ptr[tile + 1] - h0;
        for (int k = threadIdx.x; k < hn; k += kThreads) {
            const long j = halo_list[h0 + k];
            TL* d = xl + 3 * (kTile + k);
            d[0] = (TL)gx(3 * j);
            d[1] = (TL)gx(3 * j + 1);
            d[2] = (TL)gx(3 * j + 2);
        }
        __syncthreads();
    }
    if (w0 < w1) {
        int cur = -1;
        TV a0 = 0, a1 = 0, a2 = 0, xi0 = 0, xi1 = 0, xi2 = 0;
        while (true) {
            const int npos = min(pos + kU, min(w1, s_sb[sq + 1]));
            int nsq = sq;
            while (nsq < kSlices - 1 && s_sb[nsq + 1] <= npos) ++nsq;
            const bool more = npos < w1;
            uint32_t Wb[kU];
            TM Hb[kU][9];
            bool ab[kU];
            if (more) issue(npos, nsq, Wb, Hb, ab);
            if (sq != cur) {
                if (cur >= 0) {
                    TL* pc = piece + (wave + cur) * 3 * 64 + lane;
                    pc[0] = (TL)a0;
                    pc[64] = (TL)a1;
                    pc[128] = (TL)a2;
                }
                cur = sq;
                const int lr = s_row[sq * 64 + lane];
                const int li = 3 * (lr >= 0 ? lr : 0);
                xi0 = (TV)xl[li];
                xi1 = (TV)xl[li + 1];
                xi2 = (TV)xl[li + 2];
                a0 = a1 = a2 = TV(0);
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const uint32_t w = Wa[u];
                const TV m = aa[u] ? TV(1) : TV(0);
                const int lc = 3 * ((w >> 16) & 0x3FFF);
                const TV x0 = m * (TV)xl[lc], x1 = m * (TV)xl[lc + 1], x2 = m * (TV)xl[lc + 2];
                const TM* h = Ha[u];
                a0 += (TV)h[0] * x0 + (TV)h[3] * x1 + (TV)h[6] * x2;
                a1 += (TV)h[1] * x0 + (TV)h[4] * x1 + (TV)h[7] * x2;
                a2 += (TV)h[2] * x0 + (TV)h[5] * x1 + (TV)h[8] * x2;
                if (aa[u] && (w & kInTile)) {
                    const int ts = w & 0xFFFF;
                    tl[3 * ts] = (TL)((TV)h[0] * xi0 + (TV)h[1] * xi1 + (TV)h[2] * xi2);
                    tl[3 * ts + 1] = (TL)((TV)h[3] * xi0 + (TV)h[4] * xi1 + (TV)h[5] * xi2);
                    tl[3 * ts + 2] = (TL)((TV)h[6] * xi0 + (TV)h[7] * xi1 + (TV)h[8] * xi2);
                }
            }
            if (!more) break;
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                Wa[u] = Wb[u];
                aa[u] = ab[u];
#pragma unroll
                for (int c = 0; c < 9; ++c) Ha[u][c] = Hb[u][c];
            }
            pos = npos;
            sq = nsq;
        }
        TL* pc = piece + (wave + cur) * 3 * 64 + lane;
        pc[0] = (TL)a0;
        pc[64] = (TL)a1;
        pc[128] = (TL)a2;
    }
    __syncthreads();
    double dot = 0.0;
    int ylr = -1;           // with kSum1: this thread's row and its new y
    TV ys0 = 0, ys1 = 0, ys2 = 0;
    if (threadIdx.x < kTile) {
        const int sq = threadIdx.x >> 6;
        const int lr = s_row[threadIdx.x];
        if (lr >= 0) {
            const int Sq = s_sb[sq], Sq1 = s_sb[sq + 1];
            TV r0 = 0, r1 = 0, r2 = 0;
            for (int wv = 0; wv < kWaves; ++wv) {
                const int lo = (int)((long)tot * wv / kWaves), hi = (int)((long)tot * (wv + 1) / kWaves);
                if ((lo > Sq ? lo : Sq) >= (hi < Sq1 ? hi : Sq1)) continue;
                const TL* pc = piece + (wv + sq) * 3 * 64 + lane;
                r0 += (TV)pc[0];
                r1 += (TV)pc[64];
                r2 += (TV)pc[128];
            }
            const int tb = (int)(ti_r >> 16), te = tb + (int)(ti_r & 0xFFFF);
            int k = tb;
            for (; k + 4 <= te; k += 4) {
                TV v[12];
#pragma unroll
                for (int e = 0; e < 12; ++e) v[e] = (TV)tl[3 * k + e];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    r0 += v[3 * e];
                    r1 += v[3 * e + 1];
                    r2 += v[3 * e + 2];
                }
            }
            for (; k < te; ++k) {
                r0 += (TV)tl[3 * k];
                r1 += (TV)tl[3 * k + 1];
                r2 += (TV)tl[3 * k + 2];
            }
            const long i = (long)tile * kTile + lr;
            // the row's own entries, exact (the LDS copy is rounded when TL is narrower)
            const TV v0 = kExactLds ? (TV)xl[3 * lr] : gx(3 * i);
            const TV v1 = kExactLds ? (TV)xl[3 * lr + 1] : gx(3 * i + 1);
            const TV v2 = kExactLds ? (TV)xl[3 * lr + 2] : gx(3 * i + 2);
            if (kFuseP) {
                p_new[3 * i] = v0;
                p_new[3 * i + 1] = v1;
                p_new[3 * i + 2] = v2;
            }
            TV* yi = y + 3 * i;
            TV o0, o1, o2;
            if (b == TV(0)) {
                o0 = a * r0;
                o1 = a * r1;
                o2 = a * r2;
            } else {
                o0 = a * r0 + b * yi[0];
                o1 = a * r1 + b * yi[1];
                o2 = a * r2 + b * yi[2];
            }
            yi[0] = o0;
            yi[1] = o1;
            yi[2] = o2;
            if (kSum1) {
                ylr = lr;
                ys0 = o0;
                ys1 = o1;
                ys2 = o2;
            }
            if (kDot) dot = (double)v0 * (double)r0 + (double)v1 * (double)r1 + (double)v2 * (double)r2;
        }
    }
    if (kDot) {
        const double sblk = block_sum(dot, dred);
        if (threadIdx.x == 0) dot_part[tile] = sblk;
    }
    if constexpr (kSum1) {
        // the tile's new rows back in row order (piece is free once every
        // row has been summed), then the level-1 sums: a level-1 node lies in
        // one 32-row level-0 domain, so in one wave of 64 rows, whose lanes
        // group by node (the leader writes the node's sum)
        __syncthreads();
        TL* const yb = piece;
        if (ylr >= 0) {
            yb[3 * ylr] = (TL)ys0;
            yb[3 * ylr + 1] = (TL)ys1;
            yb[3 * ylr + 2] = (TL)ys2;
        }
        __syncthreads();
        if (threadIdx.x < kTile) {  // waves 0 .. kTile/64 - 1, whole
            const long u = (long)tile * kTile + threadIdx.x;
            const bool real = u < n_rows;
            const TV v0 = real ? (TV)yb[3 * threadIdx.x] : TV(0);
            const TV v1 = real ? (TV)yb[3 * threadIdx.x + 1] : TV(0);
            const TV v2 = real ? (TV)yb[3 * threadIdx.x + 2] : TV(0);
            int node = real ? l1.map1[u] : -1;
            bool pending = node >= 0;
            const int lane = threadIdx.x & 63;
            while (true) {
                const uint64_t act = __ballot(pending);
                if (act == 0) break;
                const int leader = __ffsll((unsigned long long)act) - 1;
                const int key = __shfl(node, leader, 64);
                const bool mine = pending && node == key;
                const TV s0 = wave_sum(mine ? v0 : TV(0));
                const TV s1 = wave_sum(mine ? v1 : TV(0));
                const TV s2 = wave_sum(mine ? v2 : TV(0));
                if (lane == leader) {
                    const long o = 3 * (long)(l1.npad + key);
                    y[o] = s0;
                    y[o + 1] = s1;
                    y[o + 2] = s2;
                }
                if (mine) pending = false;
            }
        }
        if (tile == 0) {
            for (int k = threadIdx.x; k < l1.pad; k += kThreads) {
                const long o = 3 * (long)(l1.npad + l1.c1 + k);
                y[o] = TV(0);
                y[o + 1] = TV(0);
                y[o + 2] = TV(0);
            }
        }
    }
}

// ---- layout build -------------------------------------------------------
__global__ void k_colrow_keys(const int* __restrict__ rows, const int* __restrict__ cols, int nnz, int kbits,
                              uint64_t* __restrict__ keys, int* __restrict__ idx)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    keys[k] = ((uint64_t)(uint32_t)cols[k] << kbits) | (uint64_t)(uint32_t)rows[k];
    idx[k] = k;
}

// up_ptr[i] = first upper block of row i; col_ptr[j] = first col-sorted entry of column j.
__global__ void k_ptrs(const int* __restrict__ rows, const uint64_t* __restrict__ skeys, int nnz, int n, int kbits,
                       int* __restrict__ up_ptr, int* __restrict__ col_ptr)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    int lo = 0, hi = nnz;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (rows[mid] < i) lo = mid + 1; else hi = mid;
    }
    up_ptr[i] = lo;
    const uint64_t key = (uint64_t)(uint32_t)i << kbits;
    lo = 0;
    hi = nnz;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (skeys[mid] < key) lo = mid + 1; else hi = mid;
    }
    col_ptr[i] = lo;
}

// Rows of more than kLongRow blocks (upper + lower, the diagonal once) are
// long: they hold nothing in the tiles and are multiplied by their own
// kernel, so a few very long rows (affine-body rows, hub vertices) cannot
// make one tile's workgroup the whole launch.
__global__ void k_long_flags(const int* __restrict__ up_ptr, const int* __restrict__ col_ptr,
                             const uint64_t* __restrict__ skeys, int n, int kbits, int* __restrict__ lflag,
                             int* __restrict__ llen)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) {
        lflag[n] = 0;
        llen[n] = 0;
        return;
    }
    // column i's entries (r, i) are sorted by r <= i: the diagonal, when
    // present, is the last one
    const int c0 = col_ptr[i], c1 = col_ptr[i + 1];
    const bool diag = c1 > c0 && (int)(skeys[c1 - 1] & ((1ull << kbits) - 1ull)) == i;
    const int L = (up_ptr[i + 1] - up_ptr[i]) + (c1 - c0) - (diag ? 1 : 0);
    const int f = L > kLongRow ? 1 : 0;
    lflag[i] = f;
    llen[i] = f ? L : 0;
}

// in-tile candidate flag per col-sorted entry (i != j, same tile, neither row long)
__global__ void k_cand(const uint64_t* __restrict__ skeys, int nnz, int kbits, const int* __restrict__ lflag,
                       int* __restrict__ cand)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nnz) return;
    const uint64_t key = skeys[s];
    const int j = (int)(key >> kbits);
    const int i = (int)(key & ((1ull << kbits) - 1ull));
    cand[s] = (i != j && i / kTile == j / kTile && !lflag[i] && !lflag[j]) ? 1 : 0;
}

// Rank -> row of the long rows and their CSR starts (one thread per row;
// thread n writes the end).
__global__ void k_long_rows(const int* __restrict__ lflag, const int* __restrict__ lrank,
                            const int* __restrict__ loff, int n, int* __restrict__ long_rows,
                            int* __restrict__ long_ptr)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) {
        long_ptr[lrank[n]] = loff[n];
        return;
    }
    if (!lflag[i]) return;
    long_rows[lrank[i]] = i;
    long_ptr[lrank[i]] = loff[i];
}

// The full rows of the long rows (CSR, one workgroup per long row, one entry
// per thread): the lower blocks (r, i), r < i, transposed, by r, then the
// upper blocks by column.  Column i's entries are sorted by r <= i, so the
// diagonal, when present, is the last of them and is taken from the upper
// blocks only.
constexpr int kLongFillThreads = 128;
__global__ __launch_bounds__(kLongFillThreads) void k_long_fill(const double* __restrict__ vals, const int* __restrict__ cols,
                            const uint64_t* __restrict__ skeys, const int* __restrict__ sidx,
                            const int* __restrict__ up_ptr, const int* __restrict__ col_ptr,
                            const int* __restrict__ long_rows, const int* __restrict__ long_ptr, int kbits,
                            int* __restrict__ long_col, double* __restrict__ lv, float* __restrict__ lv32)
{
    const int k = blockIdx.x;
    const int i = long_rows[k];
    const long e0 = long_ptr[k];
    const int c0 = col_ptr[i], c1 = col_ptr[i + 1];
    const bool diag = c1 > c0 && (int)(skeys[c1 - 1] & ((1ull << kbits) - 1ull)) == i;
    const int nl = c1 - c0 - (diag ? 1 : 0);
    const int nu = up_ptr[i + 1] - up_ptr[i];
    for (int t = threadIdx.x; t < nl + nu; t += kLongFillThreads) {
        const bool tr = t < nl;
        const int c = tr ? (int)(skeys[c0 + t] & ((1ull << kbits) - 1ull)) : cols[up_ptr[i] + t - nl];
        const int b = tr ? sidx[c0 + t] : up_ptr[i] + t - nl;
        const long e = e0 + t;
        long_col[e] = c;
        const double* h = vals + 9L * b;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const double v = tr ? h[r * 3 + q] : h[q * 3 + r];
                lv[9 * e + q * 3 + r] = v;
                if (lv32) lv32[9 * e + q * 3 + r] = (float)v;
            }
    }
}

// y_i += a (A x)_i for every long row i, one workgroup per row (fixed-order
// block reduction); the tiled kernel has already written b y_i there.  The
// x^T(A x) partial of row i goes to dot_part[k].
template <typename TM, typename TV, bool kDot>
__global__ __launch_bounds__(256) void k_spmv_long(const int* __restrict__ long_rows, const int* __restrict__ long_ptr,
                                                   const int* __restrict__ long_col, const TM* __restrict__ lv, TV a,
                                                   const TV* __restrict__ x, TV* __restrict__ y,
                                                   double* __restrict__ dot_part, const int* __restrict__ done)
{
    __shared__ TV red[16];
    if (done != nullptr && *done) return;
    const int k = blockIdx.x;
    const int i = long_rows[k];
    TV s0 = 0, s1 = 0, s2 = 0;
    for (int e = long_ptr[k] + threadIdx.x; e < long_ptr[k + 1]; e += 256) {
        const TM* h = lv + 9L * e;
        const long j = long_col[e];
        const TV x0 = x[3 * j], x1 = x[3 * j + 1], x2 = x[3 * j + 2];
        s0 += (TV)h[0] * x0 + (TV)h[3] * x1 + (TV)h[6] * x2;
        s1 += (TV)h[1] * x0 + (TV)h[4] * x1 + (TV)h[7] * x2;
        s2 += (TV)h[2] * x0 + (TV)h[5] * x1 + (TV)h[8] * x2;
    }
    s0 = block_sum(s0, red);
    s1 = block_sum(s1, red);
    s2 = block_sum(s2, red);
    if (threadIdx.x == 0) {
        TV* yi = y + 3L * i;
        yi[0] += a * s0;
        yi[1] += a * s1;
        yi[2] += a * s2;
        if (kDot)
            dot_part[k] = (double)x[3L * i] * (double)s0 + (double)x[3L * i + 1] * (double)s1 +
                          (double)x[3L * i + 2] * (double)s2;
    }
}

// Per column j: the upper word of every block (i, j), the in-tile slot range of
// column j and the count of lower (cross-tile) entries row j stores.
// A candidate whose rank in its tile reaches kTcap is stored as cross-tile.
__global__ void k_col_words(const uint64_t* __restrict__ skeys, const int* __restrict__ sidx,
                            const int* __restrict__ cscan, const int* __restrict__ col_ptr, int n, int kbits,
                            const int* __restrict__ lflag, uint32_t* __restrict__ up_word,
                            uint32_t* __restrict__ tinfo, int* __restrict__ low_cnt)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int tile0 = (j / kTile) * kTile;
    const int rank_base = cscan[col_ptr[tile0]];
    int tb = -1, tcount = 0, low = 0;
    for (int s = col_ptr[j]; s < col_ptr[j + 1]; ++s) {
        const int i = (int)(skeys[s] & ((1ull << kbits) - 1ull));
        const int b = sidx[s];
        if (i == j) {
            up_word[b] = (uint32_t)j;
            continue;
        }
        const bool same = i / kTile == j / kTile;
        const int rank = cscan[s] - rank_base;
        if (same && !lflag[i] && !lflag[j] && rank < kTcap) {
            up_word[b] = kInTile | ((uint32_t)(j - tile0) << 16) | (uint32_t)rank;
            if (tb < 0) tb = rank;
            ++tcount;
        } else {
            // row i holds H (unless long), row j holds H^T (unless long)
            up_word[b] = (uint32_t)j;
            if (!lflag[j]) ++low;
        }
    }
    tinfo[j] = ((uint32_t)(tb < 0 ? 0 : tb) << 16) | (uint32_t)tcount;
    low_cnt[j] = low;
}

__global__ void k_row_keys(const int* __restrict__ up_ptr, const int* __restrict__ low_cnt,
                           const int* __restrict__ lflag, int n, int* __restrict__ len, uint64_t* __restrict__ key,
                           int* __restrict__ iota)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int L = lflag[i] ? 0 : (up_ptr[i + 1] - up_ptr[i]) + low_cnt[i];
    len[i] = L;
    key[i] = ((uint64_t)(i / kTile) << 16) | (uint64_t)(0xFFFFu - (uint32_t)(L < 0xFFFF ? L : 0xFFFF));
    iota[i] = i;
}

// perm/plen over all tile positions (rows past n are padding: -1 / 0)
__global__ void k_perm_fill(const int* __restrict__ sorted, const int* __restrict__ len, int n, int n_pos,
                            int* __restrict__ perm, int* __restrict__ plen)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pos) return;
    const int i = p < n ? sorted[p] : -1;
    perm[p] = i;
    plen[p] = i >= 0 ? len[i] : 0;
}

__global__ void k_slice_w(const int* __restrict__ plen, int n_slices, int* __restrict__ slice_w)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > n_slices) return;
    slice_w[q] = q == n_slices ? 0 : plen[q * 64];  // longest row of the slice (sorted in its tile)
}

// The 64 lanes of a wave each need one 9-double block (b < 0: none): they
// load the 576 words together, 64 consecutive words per instruction (the
// blocks of ~7 lanes), through the wave's LDS buffer; afterwards lane l finds
// its block at buf[9 l].
__device__ __forceinline__ void wave_gather9(const double* __restrict__ vals, int b, int lane, double* __restrict__ buf)
{
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int w = 64 * t + lane;
        const int sl = w / 9, e = w - 9 * (w / 9);
        const int bs = __shfl(b, sl, 64);
        buf[w] = bs >= 0 ? vals[9L * bs + e] : 0.0;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// slot entries of a block given column major (tr: stored transposed)
__device__ __forceinline__ void put9(const double* __restrict__ s, bool tr, long slot, int lane,
                                     double* __restrict__ sv, float* __restrict__ sv32)
{
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double v = tr ? s[r * 3 + c] : s[c * 3 + r];
            sv[(slot * 9 + c * 3 + r) * 64 + lane] = v;
            if (sv32) sv32[(slot * 9 + c * 3 + r) * 64 + lane] = (float)v;
        }
}

// One thread per tile position: the row's upper blocks (diagonal first, by
// column), then its cross-tile lower entries (by row), then zero padding.
// The lanes of a wave walk their rows in step (a block per lane per step,
// loaded together by wave_gather9); the loops are per wave.
__global__ __launch_bounds__(256) void k_pack(const double* __restrict__ vals, const uint32_t* __restrict__ up_word,
                                              const uint64_t* __restrict__ skeys, const int* __restrict__ sidx,
                                              const int* __restrict__ up_ptr, const int* __restrict__ col_ptr,
                                              const int* __restrict__ perm, const int* __restrict__ slice_ptr,
                                              const int* __restrict__ slice_w, const int* __restrict__ lflag,
                                              int n_pos, int kbits, double* __restrict__ sv,
                                              float* __restrict__ sv32, uint32_t* __restrict__ scw)
{
    __shared__ double sbuf[4][576];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = p < n_pos;  // every lane stays for the wave's exchanges
    const int q = p >> 6, lane = p & 63;
    double* const buf = sbuf[threadIdx.x >> 6];
    const int i = valid ? perm[p] : -1;
    const long S = valid ? slice_ptr[q] : 0;
    const int W = valid ? slice_w[q] : 0;
    const bool held = i >= 0 && !lflag[i];  // a long row stores nothing here
    const int u0 = held ? up_ptr[i] : 0, nu = held ? up_ptr[i + 1] - u0 : 0;
    const int l0 = held ? col_ptr[i] : 0, nl = held ? col_ptr[i + 1] - l0 : 0;
    int k = 0;
    const int mu = wave_max(nu);
    for (int t = 0; t < mu; ++t) {
        const int b = t < nu ? u0 + t : -1;
        wave_gather9(vals, b, lane, buf);
        if (b >= 0) {
            scw[(S + k) * 64 + lane] = up_word[b];
            put9(buf + 9 * lane, false, S + k, lane, sv, sv32);
            ++k;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // buf is reused
    }
    const int ml = wave_max(nl);
    for (int t = 0; t < ml; ++t) {
        int b = -1, r = 0;
        if (t < nl) {
            const int e = l0 + t;
            r = (int)(skeys[e] & ((1ull << kbits) - 1ull));
            b = sidx[e];
            if (r == i || (up_word[b] & kInTile)) b = -1;
        }
        wave_gather9(vals, b, lane, buf);
        if (b >= 0) {
            scw[(S + k) * 64 + lane] = (uint32_t)r;
            put9(buf + 9 * lane, true, S + k, lane, sv, sv32);
            ++k;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    if (!valid) return;
    for (; k < W; ++k) {
        scw[(S + k) * 64 + lane] = (uint32_t)(i >= 0 ? i : 0);
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            sv[((S + k) * 9 + c) * 64 + lane] = 0.0;
            if (sv32) sv32[((S + k) * 9 + c) * 64 + lane] = 0.0f;
        }
    }
}

// ---- halo build: one thread per (slice, lane) walks that lane's slots.
// A word that is not an in-tile transposed slot names a global column; it is
// external when the column lies outside the slot's tile.  Padding rows are
// skipped (their words are rewritten to local index 0 and never used).
__global__ void k_ext_count(const uint32_t* __restrict__ cw, const int* __restrict__ slice_ptr,
                            const int* __restrict__ perm, int n_slices, int* __restrict__ cnt)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_slices * 64) return;
    const int q = t >> 6, lane = t & 63;
    int c = 0;
    if (perm[t] >= 0) {
        const int tile = q / kSlices;
        for (long s = slice_ptr[q]; s < slice_ptr[q + 1]; ++s) {
            const uint32_t w = cw[s * 64 + lane];
            if (!(w & kInTile) && (int)w / kTile != tile) ++c;
        }
    }
    cnt[t] = c;
}

__global__ void k_ext_emit(const uint32_t* __restrict__ cw, const int* __restrict__ slice_ptr,
                           const int* __restrict__ perm, const int* __restrict__ off, int n_slices,
                           uint64_t* __restrict__ keys)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_slices * 64 || perm[t] < 0) return;
    const int q = t >> 6, lane = t & 63;
    const int tile = q / kSlices;
    int o = off[t];
    for (long s = slice_ptr[q]; s < slice_ptr[q + 1]; ++s) {
        const uint32_t w = cw[s * 64 + lane];
        if (!(w & kInTile) && (int)w / kTile != tile) keys[o++] = ((uint64_t)tile << 32) | (uint64_t)w;
    }
}

// first occurrence of each (tile, vertex) in the sorted keys
__global__ void k_uniq_flag(const uint64_t* __restrict__ k, int n, int* __restrict__ flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || k[i] != k[i - 1]) ? 1 : 0;
}

// halo_list[uid] = vertex for the first occurrence (inc = inclusive scan of the flags)
__global__ void k_halo_list(const uint64_t* __restrict__ k, const int* __restrict__ flag,
                            const int* __restrict__ inc, int n, int* __restrict__ halo_list)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    halo_list[inc[i] - 1] = (int)(uint32_t)k[i];
}

// halo_ptr[t] = uid of the first key of tile t (keys of a tile are contiguous)
__global__ void k_halo_ptr(const uint64_t* __restrict__ k, const int* __restrict__ inc, int n, int n_tiles,
                           int* __restrict__ halo_ptr)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > n_tiles) return;
    const uint64_t key = (uint64_t)t << 32;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (k[mid] < key) lo = mid + 1; else hi = mid;
    }
    halo_ptr[t] = lo < n ? inc[lo] - 1 : (n > 0 ? inc[n - 1] : 0);
}

// stats[0] = largest halo of a tile, stats[1] = in-tile transposed slots a tile uses
__global__ void k_tile_stats(const int* __restrict__ halo_ptr, const uint32_t* __restrict__ tinfo, int n_tiles, int n,
                             int* __restrict__ stats)
{
    // one atomic per wave and statistic (a single address takes them all)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int h = wave_max(i < n_tiles ? halo_ptr[i + 1] - halo_ptr[i] : 0);
    const int t = wave_max(i < n ? (int)(tinfo[i] >> 16) + (int)(tinfo[i] & 0xFFFF) : 0);
    if ((threadIdx.x & 63) == 0) {
        if (h > 0) atomicMax(&stats[0], h);
        if (t > 0) atomicMax(&stats[1], t);
    }
}

// rewrite the words as local indices: own rows (column - tile start), halo
// vertices (kTile + rank in the tile's ascending halo list)
__global__ void k_localize(uint32_t* __restrict__ cw, const int* __restrict__ slice_ptr,
                           const int* __restrict__ perm, const int* __restrict__ halo_ptr,
                           const int* __restrict__ halo_list, int n_slices)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_slices * 64) return;
    const int q = t >> 6, lane = t & 63;
    const int tile = q / kSlices;
    const bool pad = perm[t] < 0;
    const int h0 = halo_ptr[tile], h1 = halo_ptr[tile + 1];
    for (long s = slice_ptr[q]; s < slice_ptr[q + 1]; ++s) {
        const uint32_t w = cw[s * 64 + lane];
        if (w & kInTile) continue;
        uint32_t lw = 0;
        if (!pad) {
            const int j = (int)w;
            if (j / kTile == tile) {
                lw = (uint32_t)(j - tile * kTile) << 16;
            } else {
                int lo = h0, hi = h1;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (halo_list[mid] < j) lo = mid + 1; else hi = mid;
                }
                lw = (uint32_t)(kTile + (lo - h0)) << 16;
            }
        }
        cw[s * 64 + lane] = lw;
    }
}

inline int g256(long n) { return (int)((n + 255) / 256); }

// Launch on stream s; with e0/e1 the events are recorded by the kernel's own
// dispatch (its start and end), not as separate stream commands.
template <typename... KArgs, typename... Args>
void go(void (*kernel)(KArgs...), int grid, int block, size_t shmem, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
        Args... args)
{
    if (e0 && e1) hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), (uint32_t)shmem, s, e0, e1, 0, args...);
    else kernel<<<grid, block, shmem, s>>>(args...);
}

}  // namespace

void FullMatrix::build(hipStream_t s, const double* vals, const int* rows, const int* cols, int nnz_up, int n,
                       bool want_f32)
{
    SGIPC_CHECK(nnz_up > 0 && n > 0, SGIPC_EMPTY, "empty matrix");
    n_rows = n;
    nnz = nnz_up;
    has_f32 = want_f32;
    int kbits = 1;
    while (kbits < 31 && (1L << kbits) < (long)n) ++kbits;
    n_tiles = (n + kTile - 1) / kTile;
    const int n_pos = n_tiles * kTile;
    n_slices = n_pos / 64;
    keys_.reserve(nnz_up);
    skeys_.reserve(nnz_up);
    idx_.reserve(nnz_up);
    sidx_.reserve(nnz_up);
    cand_.reserve(nnz_up + 1);
    cscan_.reserve(nnz_up + 1);
    up_word_.reserve(nnz_up);
    up_ptr_.reserve(n + 1);
    col_ptr_.reserve(n + 1);
    low_cnt_.reserve(n);
    len_.reserve(n);
    rkey_.reserve(n);
    rkey_sorted_.reserve(n);
    iota_.reserve(n);
    sorted_.reserve(n);
    perm.reserve(n_pos);
    plen.reserve(n_pos);
    tinfo.reserve(n);
    slice_w.reserve(n_slices + 1);
    slice_ptr.reserve(n_slices + 1);
    lflag_.reserve(n + 1);
    llen_.reserve(n + 1);
    lrank_.reserve(n + 1);
    loff_.reserve(n + 1);

    int tbits = 1;
    while (tbits < 40 && (1L << tbits) < (long)n_tiles) ++tbits;
    size_t need = 0, need2 = 0, need3 = 0, need4 = 0;
    SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, need, keys_.get(), skeys_.get(), idx_.get(), sidx_.get(),
                                        (size_t)nnz_up, 0, 2 * kbits, s));
    SGIPC_HIP(rocprim::exclusive_scan(nullptr, need2, cand_.get(), cscan_.get(), 0, (size_t)nnz_up + 1,
                                      rocprim::plus<int>(), s));
    SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, need3, rkey_.get(), rkey_sorted_.get(), iota_.get(), sorted_.get(),
                                        (size_t)n, 0, 16 + tbits, s));
    SGIPC_HIP(rocprim::exclusive_scan(nullptr, need4, slice_w.get(), slice_ptr.get(), 0, (size_t)n_slices + 1,
                                      rocprim::plus<int>(), s));
    size_t need5 = 0;
    SGIPC_HIP(rocprim::exclusive_scan(nullptr, need5, lflag_.get(), lrank_.get(), 0, (size_t)n + 1,
                                      rocprim::plus<int>(), s));
    size_t nt = need;
    if (need5 > nt) nt = need5;
    if (need2 > nt) nt = need2;
    if (need3 > nt) nt = need3;
    if (need4 > nt) nt = need4;
    temp_.reserve(nt);

    // column-major order of the upper blocks
    k_colrow_keys<<<g256(nnz_up), 256, 0, s>>>(rows, cols, nnz_up, kbits, keys_.get(), idx_.get());
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(rocprim::radix_sort_pairs(temp_.get(), need, keys_.get(), skeys_.get(), idx_.get(), sidx_.get(),
                                        (size_t)nnz_up, 0, 2 * kbits, s));
    k_ptrs<<<g256(n + 1), 256, 0, s>>>(rows, skeys_.get(), nnz_up, n, kbits, up_ptr_.get(), col_ptr_.get());
    SGIPC_LAUNCH_CHECK();
    // long rows: flags, ranks and the offsets of their full rows
    k_long_flags<<<g256(n + 1), 256, 0, s>>>(up_ptr_.get(), col_ptr_.get(), skeys_.get(), n, kbits, lflag_.get(),
                                             llen_.get());
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(rocprim::exclusive_scan(temp_.get(), need5, lflag_.get(), lrank_.get(), 0, (size_t)n + 1,
                                      rocprim::plus<int>(), s));
    SGIPC_HIP(rocprim::exclusive_scan(temp_.get(), need5, llen_.get(), loff_.get(), 0, (size_t)n + 1,
                                      rocprim::plus<int>(), s));
    // in-tile ranks: exclusive scan of the candidate flags in column order
    k_cand<<<g256(nnz_up), 256, 0, s>>>(skeys_.get(), nnz_up, kbits, lflag_.get(), cand_.get());
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(hipMemsetAsync(cand_.get() + nnz_up, 0, sizeof(int), s));
    SGIPC_HIP(rocprim::exclusive_scan(temp_.get(), need2, cand_.get(), cscan_.get(), 0, (size_t)nnz_up + 1,
                                      rocprim::plus<int>(), s));
    k_col_words<<<g256(n), 256, 0, s>>>(skeys_.get(), sidx_.get(), cscan_.get(), col_ptr_.get(), n, kbits,
                                        lflag_.get(), up_word_.get(), tinfo.get(), low_cnt_.get());
    SGIPC_LAUNCH_CHECK();
    // rows sorted by entry count inside each tile
    k_row_keys<<<g256(n), 256, 0, s>>>(up_ptr_.get(), low_cnt_.get(), lflag_.get(), n, len_.get(), rkey_.get(),
                                       iota_.get());
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(rocprim::radix_sort_pairs(temp_.get(), need3, rkey_.get(), rkey_sorted_.get(), iota_.get(),
                                        sorted_.get(), (size_t)n, 0, 16 + tbits, s));
    k_perm_fill<<<g256(n_pos), 256, 0, s>>>(sorted_.get(), len_.get(), n, n_pos, perm.get(), plen.get());
    SGIPC_LAUNCH_CHECK();
    k_slice_w<<<g256(n_slices + 1), 256, 0, s>>>(plen.get(), n_slices, slice_w.get());
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(rocprim::exclusive_scan(temp_.get(), need4, slice_w.get(), slice_ptr.get(), 0, (size_t)n_slices + 1,
                                      rocprim::plus<int>(), s));
    // one read back for the slot count and the long rows' counts
    int hc[3] = {0, 0, 0};
    SGIPC_HIP(hipMemcpyAsync(&hc[0], slice_ptr.get() + n_slices, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipMemcpyAsync(&hc[1], lrank_.get() + n, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipMemcpyAsync(&hc[2], loff_.get() + n, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    SGIPC_CHECK(hc[0] >= 0, SGIPC_CAPACITY, "SELL slot count overflows int");
    SGIPC_CHECK(hc[2] >= 0, SGIPC_CAPACITY, "long-row entry count overflows int");
    n_slots = hc[0];
    n_long = hc[1];
    {
        const int ne = hc[2] > 0 ? hc[2] : 1;
        long_rows.reserve(n_long > 0 ? n_long : 1);
        long_ptr.reserve((size_t)n_long + 1);
        long_col.reserve(ne);
        long_vals.reserve((size_t)ne * 9);
        if (want_f32) long_vals32.reserve((size_t)ne * 9);
        k_long_rows<<<g256(n + 1), 256, 0, s>>>(lflag_.get(), lrank_.get(), loff_.get(), n, long_rows.get(),
                                                long_ptr.get());
        SGIPC_LAUNCH_CHECK();
        if (n_long > 0) {
            k_long_fill<<<n_long, kLongFillThreads, 0, s>>>(vals, cols, skeys_.get(), sidx_.get(), up_ptr_.get(),
                                                            col_ptr_.get(), long_rows.get(), long_ptr.get(), kbits,
                                                            long_col.get(), long_vals.get(),
                                                            want_f32 ? long_vals32.get() : nullptr);
            SGIPC_LAUNCH_CHECK();
        }
    }
    sell_vals.reserve((size_t)n_slots * 64 * 9);
    sell_cols.reserve((size_t)n_slots * 64);
    if (want_f32) sell_vals32.reserve((size_t)n_slots * 64 * 9);
    k_pack<<<g256(n_pos), 256, 0, s>>>(vals, up_word_.get(), skeys_.get(), sidx_.get(), up_ptr_.get(), col_ptr_.get(),
                                       perm.get(), slice_ptr.get(), slice_w.get(), lflag_.get(), n_pos, kbits,
                                       sell_vals.get(), want_f32 ? sell_vals32.get() : nullptr, sell_cols.get());
    SGIPC_LAUNCH_CHECK();
    build_halo(s, tbits);
}

void FullMatrix::build_halo(hipStream_t s, int tbits)
{
    use_halo = false;
    const int nl = n_slices * 64;
    hcnt_.reserve(nl + 1);
    hoff_.reserve(nl + 1);
    halo_ptr.reserve(n_tiles + 1);
    stats_.reserve(2);
    size_t need = 0;
    SGIPC_HIP(rocprim::exclusive_scan(nullptr, need, hcnt_.get(), hoff_.get(), 0, (size_t)nl + 1, rocprim::plus<int>(),
                                      s));
    temp_.reserve(need);
    k_ext_count<<<g256(nl), 256, 0, s>>>(sell_cols.get(), slice_ptr.get(), perm.get(), n_slices, hcnt_.get());
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(hipMemsetAsync(hcnt_.get() + nl, 0, sizeof(int), s));
    SGIPC_HIP(rocprim::exclusive_scan(temp_.get(), need, hcnt_.get(), hoff_.get(), 0, (size_t)nl + 1,
                                      rocprim::plus<int>(), s));
    int n_ext = 0;
    SGIPC_HIP(hipMemcpyAsync(&n_ext, hoff_.get() + nl, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    SGIPC_CHECK(n_ext >= 0, SGIPC_CAPACITY, "halo entry count overflows int");
    if (n_ext > 0) {
        hkeys_.reserve(n_ext);
        hkeys_sorted_.reserve(n_ext);
        hflag_.reserve(n_ext);
        hinc_.reserve(n_ext);
        halo_list.reserve(n_ext);
        size_t n1 = 0, n2 = 0;
        SGIPC_HIP(rocprim::radix_sort_keys(nullptr, n1, hkeys_.get(), hkeys_sorted_.get(), (size_t)n_ext, 0, 32 + tbits,
                                           s));
        SGIPC_HIP(rocprim::inclusive_scan(nullptr, n2, hflag_.get(), hinc_.get(), (size_t)n_ext, rocprim::plus<int>(),
                                          s));
        temp_.reserve(n1 > n2 ? n1 : n2);
        k_ext_emit<<<g256(nl), 256, 0, s>>>(sell_cols.get(), slice_ptr.get(), perm.get(), hoff_.get(), n_slices,
                                            hkeys_.get());
        SGIPC_LAUNCH_CHECK();
        SGIPC_HIP(rocprim::radix_sort_keys(temp_.get(), n1, hkeys_.get(), hkeys_sorted_.get(), (size_t)n_ext, 0,
                                           32 + tbits, s));
        k_uniq_flag<<<g256(n_ext), 256, 0, s>>>(hkeys_sorted_.get(), n_ext, hflag_.get());
        SGIPC_LAUNCH_CHECK();
        SGIPC_HIP(rocprim::inclusive_scan(temp_.get(), n2, hflag_.get(), hinc_.get(), (size_t)n_ext,
                                          rocprim::plus<int>(), s));
        k_halo_list<<<g256(n_ext), 256, 0, s>>>(hkeys_sorted_.get(), hflag_.get(), hinc_.get(), n_ext,
                                                halo_list.get());
        SGIPC_LAUNCH_CHECK();
    } else {
        halo_list.reserve(1);
    }
    k_halo_ptr<<<g256(n_tiles + 1), 256, 0, s>>>(hkeys_sorted_.get(), hinc_.get(), n_ext, n_tiles, halo_ptr.get());
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(hipMemsetAsync(stats_.get(), 0, 2 * sizeof(int), s));
    k_tile_stats<<<g256(n_tiles > n_rows ? n_tiles : n_rows), 256, 0, s>>>(halo_ptr.get(), tinfo.get(), n_tiles, n_rows,
                                                                             stats_.get());
    SGIPC_LAUNCH_CHECK();
    int st[2] = {0, 0};
    SGIPC_HIP(hipMemcpyAsync(st, stats_.get(), sizeof(st), hipMemcpyDeviceToHost, s));
    std::vector<int> sp((size_t)n_slices + 1);
    SGIPC_HIP(hipMemcpyAsync(sp.data(), slice_ptr.get(), sizeof(int) * sp.size(), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    halo_max = st[0];
    tl_max = st[1] > 0 ? st[1] : 1;
    // the FP64-vector kernels have the largest LDS; use the halo form only when
    // it fits beside the largest static LDS among them
    if (halo_dyn_bytes(sizeof(double), kThreads) + halo_static_lds() > (size_t)kLdsBytes) return;
    const char* env = std::getenv("SGIPC_SPMV_GATHER");
    if (env && env[0] && env[0] != '0') return;  // keep the global-gather form (tests of both forms)
    k_localize<<<g256(nl), 256, 0, s>>>(sell_cols.get(), slice_ptr.get(), perm.get(), halo_ptr.get(), halo_list.get(),
                                        n_slices);
    SGIPC_LAUNCH_CHECK();
    // heaviest tiles first inside each XCD's contiguous range of the launch order
    std::vector<int> ord((size_t)n_tiles);
    const int q = n_tiles / 8, rem = n_tiles % 8;
    for (int xc = 0; xc < 8; ++xc) {
        const int b0 = xc * q + (xc < rem ? xc : rem), cnt = q + (xc < rem ? 1 : 0);
        for (int k = 0; k < cnt; ++k) ord[(size_t)b0 + k] = b0 + k;
        std::stable_sort(ord.begin() + b0, ord.begin() + b0 + cnt, [&](int u, int v) {
            return sp[(size_t)(u + 1) * kSlices] - sp[(size_t)u * kSlices] >
                   sp[(size_t)(v + 1) * kSlices] - sp[(size_t)v * kSlices];
        });
    }
    order.reserve(n_tiles);
    SGIPC_HIP(hipMemcpyAsync(order.get(), ord.data(), sizeof(int) * ord.size(), hipMemcpyHostToDevice, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    use_halo = true;
}

size_t FullMatrix::halo_dyn_bytes(size_t tv_size, int threads) const
{
    const size_t pieces = (size_t)(threads / 64 + kSlices - 1);
    return (3 * (size_t)tl_max + pieces * 3 * 64 + 3 * ((size_t)kTile + (size_t)halo_max)) * tv_size;
}

size_t FullMatrix::halo_static_lds()
{
    static size_t cached = 0;
    if (cached == 0) {
        const void* fns[] = {reinterpret_cast<const void*>(&k_spmv_halo<double, double, false, false>),
                             reinterpret_cast<const void*>(&k_spmv_halo<double, double, true, false>),
                             reinterpret_cast<const void*>(&k_spmv_halo<double, double, false, true>),
                             reinterpret_cast<const void*>(&k_spmv_halo<double, double, true, true>),
                             reinterpret_cast<const void*>(&k_spmv_halo<float, double, true, false>),
                             reinterpret_cast<const void*>(&k_spmv_halo<float, double, true, true>),
                             reinterpret_cast<const void*>(&k_spmv_halo<double, double, false, false, true>)};
        size_t m = 0;
        for (const void* f : fns) {
            hipFuncAttributes fa;
            SGIPC_HIP(hipFuncGetAttributes(&fa, f));
            if (fa.sharedSizeBytes > m) m = fa.sharedSizeBytes;
        }
        cached = m;
    }
    return cached;
}

int FullMatrix::partials() const { return n_tiles + n_long; }

long FullMatrix::stored_blocks(hipStream_t s) const
{
    std::vector<int> h((size_t)n_tiles * kTile);
    SGIPC_HIP(hipMemcpyAsync(h.data(), plen.get(), sizeof(int) * h.size(), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    long t = 0;
    for (int v : h) t += v;
    return t;
}

template <typename TM, typename TV>
int FullMatrix::launch(hipStream_t s, const TM* vals, TV a, const TV* x, TV b, TV* y, double* dot_part,
                       const int* done, const TV* z, TV* p_new, const double* rzn, const double* rzo,
                       int n_rz, const Level1Sums* l1) const
{
    SGIPC_CHECK(n_tiles > 0, SGIPC_EMPTY, "matrix was not built");
    const int g = n_tiles;
    SGIPC_CHECK(l1 == nullptr || (use_halo && !dot_part && !p_new), SGIPC_BAD_ARG,
                "level-1 sums ride on the plain halo-form SpMV only");
    const Level1Sums l1v = l1 ? *l1 : Level1Sums{};
    if (use_halo) {
        constexpr int nt = HaloCfg<TM, TV>::NT;
        const size_t dyn = halo_dyn_bytes(sizeof(typename HaloCfg<TM, TV>::TL), nt);
#define SGIPC_HALO(DOT, FUSE, SUM1)                                                                                 \
    do {                                                                                                            \
        const void* fn = reinterpret_cast<const void*>(&k_spmv_halo<TM, TV, DOT, FUSE, SUM1>);                      \
        static int max_dyn = -1;                                                                                    \
        if (max_dyn < 0) {                                                                                          \
            hipFuncAttributes fa;                                                                                   \
            SGIPC_HIP(hipFuncGetAttributes(&fa, fn));                                                               \
            const int m = kLdsBytes - (int)fa.sharedSizeBytes;                                                      \
            SGIPC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, m));                      \
            max_dyn = m;                                                                                            \
        }                                                                                                           \
        SGIPC_CHECK(dyn <= (size_t)max_dyn, SGIPC_CAPACITY, "halo SpMV needs more LDS than a CU has");              \
        go(&k_spmv_halo<TM, TV, DOT, FUSE, SUM1>, g, nt, dyn, s, time_start, time_stop, vals,                      \
           (const uint32_t*)sell_cols.get(), (const int*)slice_ptr.get(), (const int*)perm.get(),                   \
           (const int*)plen.get(), (const uint32_t*)tinfo.get(), (const int*)halo_ptr.get(),                        \
           (const int*)halo_list.get(), (const int*)order.get(), tl_max, a, x, b, y, dot_part, done, z, p_new, rzn, \
           rzo, n_rz, n_rows, l1v);                                                                                 \
    } while (0)
        if (l1) {
            SGIPC_HALO(false, false, true);
        } else if (p_new) {
            if (dot_part) SGIPC_HALO(true, true, false);
            else SGIPC_HALO(false, true, false);
        } else {
            if (dot_part) SGIPC_HALO(true, false, false);
            else SGIPC_HALO(false, false, false);
        }
#undef SGIPC_HALO
        SGIPC_LAUNCH_CHECK();
        launch_long<TM, TV>(s, a, p_new ? p_new : x, y, dot_part, done);
        return dot_part ? g + n_long : 0;
    }
#define SGIPC_TSYM(DOT, FUSE)                                                                                   \
    go(&k_spmv_tsym<TM, TV, DOT, FUSE>, g, kThreads, 0, s, time_start, time_stop, vals,                            \
       (const uint32_t*)sell_cols.get(),                                                                            \
       (const int*)slice_ptr.get(), (const int*)slice_w.get(), (const int*)perm.get(), (const int*)plen.get(),       \
       (const uint32_t*)tinfo.get(), a, x, b, y, dot_part, done, z, p_new, rzn, rzo, n_rz, n_rows)
    if (p_new) {
        if (dot_part) SGIPC_TSYM(true, true);
        else SGIPC_TSYM(false, true);
    } else {
        if (dot_part) SGIPC_TSYM(true, false);
        else SGIPC_TSYM(false, false);
    }
#undef SGIPC_TSYM
    SGIPC_LAUNCH_CHECK();
    launch_long<TM, TV>(s, a, p_new ? p_new : x, y, dot_part, done);
    return dot_part ? g + n_long : 0;
}

template <typename TM, typename TV>
void FullMatrix::launch_long(hipStream_t s, TV a, const TV* x, TV* y, double* dot_part, const int* done) const
{
    if (n_long <= 0) return;
    // FP64 vectors read the long rows' FP64 values also when the tiles' are FP32
    using TLV = typename std::conditional<sizeof(TV) == 8, double, TM>::type;
    const TLV* lv;
    if constexpr (sizeof(TLV) == 8) lv = long_vals.get();
    else lv = long_vals32.get();
    double* dp = dot_part ? dot_part + n_tiles : nullptr;
    if (dot_part)
        go(&k_spmv_long<TLV, TV, true>, n_long, 256, 0, s, nullptr, nullptr, (const int*)long_rows.get(),
           (const int*)long_ptr.get(), (const int*)long_col.get(), lv, a, x, y, dp, done);
    else
        go(&k_spmv_long<TLV, TV, false>, n_long, 256, 0, s, nullptr, nullptr, (const int*)long_rows.get(),
           (const int*)long_ptr.get(), (const int*)long_col.get(), lv, a, x, y, dp, done);
    SGIPC_LAUNCH_CHECK();
}

template <typename TV>
int FullMatrix::apply(hipStream_t s, TV a, const TV* x, TV b, TV* y, double* dot_part, const int* done) const
{
    if constexpr (sizeof(TV) == 8) {
        return launch<double, double>(s, sell_vals.get(), a, x, b, y, dot_part, done, nullptr, nullptr, nullptr,
                                      nullptr, 0);
    } else {
        SGIPC_CHECK(has_f32, SGIPC_BAD_ARG, "FP32 copy of the matrix was not built");
        return launch<float, float>(s, sell_vals32.get(), a, x, b, y, dot_part, done, nullptr, nullptr, nullptr,
                                    nullptr, 0);
    }
}

template <typename TV>
bool FullMatrix::apply_level1_sums(hipStream_t s, TV a, const TV* x, TV b, TV* y, const int* done,
                                   const Level1Sums& l1) const
{
    if (!use_halo || n_tiles <= 0 || n_long > 0) return false;  // the epilogue needs every row in the tiles
    SGIPC_CHECK(l1.map1 != nullptr && l1.npad >= n_rows && l1.c1 >= 0 && l1.pad >= 0, SGIPC_BAD_ARG,
                "level-1 sums: bad slot layout");
    if constexpr (sizeof(TV) == 8) {
        launch<double, double>(s, sell_vals.get(), a, x, b, y, nullptr, done, nullptr, nullptr, nullptr, nullptr, 0,
                               &l1);
    } else {
        SGIPC_CHECK(has_f32, SGIPC_BAD_ARG, "FP32 copy of the matrix was not built");
        launch<float, float>(s, sell_vals32.get(), a, x, b, y, nullptr, done, nullptr, nullptr, nullptr, nullptr, 0,
                             &l1);
    }
    return true;
}

template bool FullMatrix::apply_level1_sums<double>(hipStream_t, double, const double*, double, double*, const int*,
                                                    const Level1Sums&) const;
template bool FullMatrix::apply_level1_sums<float>(hipStream_t, float, const float*, float, float*, const int*,
                                                   const Level1Sums&) const;

int FullMatrix::apply_f32_values(hipStream_t s, double a, const double* x, double b, double* y, double* dot_part,
                                 const int* done) const
{
    SGIPC_CHECK(has_f32, SGIPC_BAD_ARG, "FP32 copy of the matrix was not built");
    return launch<float, double>(s, sell_vals32.get(), a, x, b, y, dot_part, done, nullptr, nullptr, nullptr,
                                 nullptr, 0);
}

template <typename TV>
int FullMatrix::apply_update_p(hipStream_t s, bool f32_values, const TV* z, const TV* p_old, TV* p_new, TV* Ap,
                               double* dot_part, const int* done, const double* rz_new, const double* rz_old,
                               int n_rz) const
{
    if constexpr (sizeof(TV) == 8) {
        if (f32_values) {
            SGIPC_CHECK(has_f32, SGIPC_BAD_ARG, "FP32 copy of the matrix was not built");
            return launch<float, double>(s, sell_vals32.get(), 1.0, p_old, 0.0, Ap, dot_part, done, z, p_new, rz_new,
                                         rz_old, n_rz);
        }
        return launch<double, double>(s, sell_vals.get(), 1.0, p_old, 0.0, Ap, dot_part, done, z, p_new, rz_new,
                                      rz_old, n_rz);
    } else {
        SGIPC_CHECK(has_f32, SGIPC_BAD_ARG, "FP32 copy of the matrix was not built");
        return launch<float, float>(s, sell_vals32.get(), 1.0f, p_old, 0.0f, Ap, dot_part, done, z, p_new, rz_new,
                                    rz_old, n_rz);
    }
}

template int FullMatrix::apply<double>(hipStream_t, double, const double*, double, double*, double*, const int*) const;
template int FullMatrix::apply<float>(hipStream_t, float, const float*, float, float*, double*, const int*) const;
template int FullMatrix::apply_update_p<double>(hipStream_t, bool, const double*, const double*, double*, double*,
                                                double*, const int*, const double*, const double*, int) const;
template int FullMatrix::apply_update_p<float>(hipStream_t, bool, const float*, const float*, float*, float*, double*,
                                               const int*, const double*, const double*, int) const;

}  // namespace sgipc
