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
#include <cstring>
#include <vector>

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

// in-tile candidate flag per col-sorted entry (i != j, same tile)
__global__ void k_cand(const uint64_t* __restrict__ skeys, int nnz, int kbits, int* __restrict__ cand)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nnz) return;
    const uint64_t key = skeys[s];
    const int j = (int)(key >> kbits);
    const int i = (int)(key & ((1ull << kbits) - 1ull));
    cand[s] = (i != j && i / kTile == j / kTile) ? 1 : 0;
}

// Per column j: the upper word of every block (i, j), the in-tile slot range of
// column j and the count of lower (cross-tile) entries row j stores.
// A candidate whose rank in its tile reaches kTcap is stored as cross-tile.
__global__ void k_col_words(const uint64_t* __restrict__ skeys, const int* __restrict__ sidx,
                            const int* __restrict__ cscan, const int* __restrict__ col_ptr, int n, int kbits,
                            uint32_t* __restrict__ up_word, uint32_t* __restrict__ tinfo, int* __restrict__ low_cnt)
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
        if (same && rank < kTcap) {
            up_word[b] = kInTile | ((uint32_t)(j - tile0) << 16) | (uint32_t)rank;
            if (tb < 0) tb = rank;
            ++tcount;
        } else {
            up_word[b] = (uint32_t)j;
            ++low;
        }
    }
    tinfo[j] = ((uint32_t)(tb < 0 ? 0 : tb) << 16) | (uint32_t)tcount;
    low_cnt[j] = low;
}

__global__ void k_row_keys(const int* __restrict__ up_ptr, const int* __restrict__ low_cnt, int n,
                           int* __restrict__ len, uint64_t* __restrict__ key, int* __restrict__ iota)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int L = (up_ptr[i + 1] - up_ptr[i]) + low_cnt[i];
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
__global__ void k_pack(const double* __restrict__ vals, const uint32_t* __restrict__ up_word,
                       const uint64_t* __restrict__ skeys, const int* __restrict__ sidx,
                       const int* __restrict__ up_ptr, const int* __restrict__ col_ptr, const int* __restrict__ perm,
                       const int* __restrict__ slice_ptr, const int* __restrict__ slice_w, int n_pos, int kbits,
                       double* __restrict__ sv, float* __restrict__ sv32, uint32_t* __restrict__ scw)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pos) return;
    const int q = p >> 6, lane = p & 63;
    const int i = perm[p];
    const long S = slice_ptr[q];
    const int W = slice_w[q];
    int k = 0;
    if (i >= 0) {
        for (int b = up_ptr[i]; b < up_ptr[i + 1]; ++b, ++k) {
            scw[(S + k) * 64 + lane] = up_word[b];
            put9(vals + 9 * (long)b, false, S + k, lane, sv, sv32);
        }
        for (int s = col_ptr[i]; s < col_ptr[i + 1]; ++s) {
            const int r = (int)(skeys[s] & ((1ull << kbits) - 1ull));
            const int b = sidx[s];
            if (r == i || (up_word[b] & kInTile)) continue;
            scw[(S + k) * 64 + lane] = (uint32_t)r;
            put9(vals + 9 * (long)b, true, S + k, lane, sv, sv32);
            ++k;
        }
    }
    for (; k < W; ++k) {
        scw[(S + k) * 64 + lane] = (uint32_t)(i >= 0 ? i : 0);
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            sv[((S + k) * 9 + c) * 64 + lane] = 0.0;
            if (sv32) sv32[((S + k) * 9 + c) * 64 + lane] = 0.0f;
        }
    }
}

inline int g256(long n) { return (int)((n + 255) / 256); }

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
    size_t nt = need;
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
    // in-tile ranks: exclusive scan of the candidate flags in column order
    k_cand<<<g256(nnz_up), 256, 0, s>>>(skeys_.get(), nnz_up, kbits, cand_.get());
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(hipMemsetAsync(cand_.get() + nnz_up, 0, sizeof(int), s));
    SGIPC_HIP(rocprim::exclusive_scan(temp_.get(), need2, cand_.get(), cscan_.get(), 0, (size_t)nnz_up + 1,
                                      rocprim::plus<int>(), s));
    k_col_words<<<g256(n), 256, 0, s>>>(skeys_.get(), sidx_.get(), cscan_.get(), col_ptr_.get(), n, kbits,
                                        up_word_.get(), tinfo.get(), low_cnt_.get());
    SGIPC_LAUNCH_CHECK();
    // rows sorted by entry count inside each tile
    k_row_keys<<<g256(n), 256, 0, s>>>(up_ptr_.get(), low_cnt_.get(), n, len_.get(), rkey_.get(), iota_.get());
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(rocprim::radix_sort_pairs(temp_.get(), need3, rkey_.get(), rkey_sorted_.get(), iota_.get(),
                                        sorted_.get(), (size_t)n, 0, 16 + tbits, s));
    k_perm_fill<<<g256(n_pos), 256, 0, s>>>(sorted_.get(), len_.get(), n, n_pos, perm.get(), plen.get());
    SGIPC_LAUNCH_CHECK();
    k_slice_w<<<g256(n_slices + 1), 256, 0, s>>>(plen.get(), n_slices, slice_w.get());
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(rocprim::exclusive_scan(temp_.get(), need4, slice_w.get(), slice_ptr.get(), 0, (size_t)n_slices + 1,
                                      rocprim::plus<int>(), s));
    int h_slots = 0;
    SGIPC_HIP(hipMemcpyAsync(&h_slots, slice_ptr.get() + n_slices, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    SGIPC_CHECK(h_slots >= 0, SGIPC_CAPACITY, "SELL slot count overflows int");
    n_slots = h_slots;
    sell_vals.reserve((size_t)n_slots * 64 * 9);
    sell_cols.reserve((size_t)n_slots * 64);
    if (want_f32) sell_vals32.reserve((size_t)n_slots * 64 * 9);
    k_pack<<<g256(n_pos), 256, 0, s>>>(vals, up_word_.get(), skeys_.get(), sidx_.get(), up_ptr_.get(), col_ptr_.get(),
                                       perm.get(), slice_ptr.get(), slice_w.get(), n_pos, kbits, sell_vals.get(),
                                       want_f32 ? sell_vals32.get() : nullptr, sell_cols.get());
    SGIPC_LAUNCH_CHECK();
}

int FullMatrix::partials() const { return n_tiles; }

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
                       int n_rz) const
{
    SGIPC_CHECK(n_tiles > 0, SGIPC_EMPTY, "matrix was not built");
    const int g = n_tiles;
#define SGIPC_TSYM(DOT, FUSE)                                                                                   \
    k_spmv_tsym<TM, TV, DOT, FUSE><<<g, kThreads, 0, s>>>(vals, sell_cols.get(), slice_ptr.get(), slice_w.get(), \
                                                          perm.get(), plen.get(), tinfo.get(), a, x, b, y,        \
                                                          dot_part, done, z, p_new, rzn, rzo, n_rz, n_rows)
    if (p_new) {
        if (dot_part) SGIPC_TSYM(true, true);
        else SGIPC_TSYM(false, true);
    } else {
        if (dot_part) SGIPC_TSYM(true, false);
        else SGIPC_TSYM(false, false);
    }
#undef SGIPC_TSYM
    SGIPC_LAUNCH_CHECK();
    return dot_part ? g : 0;
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
