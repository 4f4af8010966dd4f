// Multilevel additive Schwarz (MAS) preconditioner for gfx950.
//
// Semantics follow StiffGIPC/MASPreconditioner.cu (GpuMas backend, identity
// slot map):
//  * hierarchy (ReorderRealtime :1785-1826, kernels :29-625, contact
//    connectivity :1132-1548): at level l the nodes are cut into domains of B
//    consecutive ids; mesh edges and contact cliques mapped to level l that
//    join two nodes of one domain are OR-ed into per-node bitmasks; each
//    domain's connected components become the level-(l+1) nodes, numbered
//    domain by domain and, inside a domain, by their lowest member;
//    off_1 = ceil(n/B)*B and off_{l+1} = off_l + ceil(c_l/B)*B;
//  * Galerkin domain matrices (PrepareHessian_bcoo :1829-2106): P_l^T A P_l
//    restricted to node pairs of one level-l domain, upper 3x3 blocks;
//  * FP64 Gauss-Jordan inverse without pivoting, zero diagonal -> 1 (:630-726);
//  * apply (preconditioning :2248-2289): restrict, per-domain dense mat-vec,
//    prolong.
//
// MI355X mapping:
//  * level-0 domain matrices are never materialised: the inverse kernel pulls
//    each domain's blocks straight from the sorted BCOO into register tiles
//    and inverts in FP64; the coarse operators come from recursive
//    coarsening: every upper block is mapped one level up (H, H^T, or
//    H + H^T on a merged diagonal), merged by the deterministic triplet
//    converter, and the domain blocks are extracted from the result;
//  * the inverse keeps a 6x6 tile of the 3Bx3B matrix per thread (256
//    threads for B=32) and broadcasts the pivot row/column through LDS, one
//    barrier per pivot (double-buffered);
//  * restriction is a deterministic wavefront group-by (no atomics): every
//    coarse node lives inside one finer domain, so its sum is formed by one
//    wave; the local solve stages a domain's FP32/FP64 inverse blocks in LDS
//    with coalesced 16-byte loads and computes each row in a fixed order;
//    prolongation fuses the r.z partial for PCG.
#include "mas.hpp"

#include <hip/hip_fp16.h>
#include "assembly.hpp"
#include "spmv_full.hpp"
#include "wave_ops.hpp"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>

namespace sgipc {

namespace {

__device__ __host__ __forceinline__ int blk_index(int B, int r, int c) { return B * r - r * (r + 1) / 2 + c; }

// ------------------------------------------------------------------ hierarchy
// OR of mask m into dst[key], combined over the lanes of the wave that share
// a key: on the coarse levels whole waves of fine nodes map to one or two
// coarse nodes, and one atomic per lane would serialise on those words.
// After a few distinct keys the remaining lanes issue their own atomics
// (keys are then spread out).  Every lane of the wave must call it.
__device__ __forceinline__ void wave_or_atomic(uint32_t* __restrict__ dst, int key, uint32_t m)
{
    const int lane = lane_id();
    uint64_t pending = __ballot(m != 0u);
    for (int round = 0; round < 4 && pending; ++round) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int k = __shfl(key, leader);
        const bool mine = ((pending >> lane) & 1ull) && key == k;
        pending &= ~__ballot(mine);
        uint32_t v = mine ? m : 0u;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v |= (uint32_t)__shfl_xor((int)v, off);
        if (lane == leader) atomicOr(dst + k, v);
    }
    if ((pending >> lane) & 1ull) atomicOr(dst + key, m);
}

template <int B>
__global__ __launch_bounds__(256) void k_mask_fine(int level, int n, const int* __restrict__ map_l,
                                                   const uint32_t* __restrict__ ns, const uint32_t* __restrict__ nn,
                                                   const uint32_t* __restrict__ nl, uint32_t* __restrict__ mask)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = u < n;
    int a = 0;
    uint32_t m = 0;
    if (act) {
        a = map_l ? map_l[u] : u;
        const uint32_t st = ns[u], cnt = nn[u];
        for (uint32_t k = 0; k < cnt; ++k) {
            int v = (int)nl[st + k];
            int b = map_l ? map_l[v] : v;
            if (a != b && a / B == b / B) m |= 1u << (b % B);
        }
    }
    if (level == 0) {
        if (act) mask[u] = m;
    } else {
        wave_or_atomic(mask, a, m);
    }
}

__device__ __forceinline__ void decode_pair(int4 pr, int coff, int n, int v[4])
{
    int x = pr.x, y = pr.y, z = pr.z, w = pr.w;
    int cnt = 4;
    if (x >= 0) {
        if (w < 0) w = -w - 1;
    } else {
        x = -x - 1;
        if (z < 0) {
            if (y < 0) {
                y = -y - 1; z = -z - 1; w = -w - 1;
            } else {
                cnt = 2;
            }
        } else if (w < 0) {
            if (y < 0) {
                y = -y - 1; w = -w - 1;
            } else {
                cnt = 3;
            }
        }
    }
    int t[4] = {x, y, z, w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int q = t[i] - coff;
        v[i] = (i < cnt && q >= 0 && q < n) ? q : -1;
    }
}

// Masks contributed by the contact pairs.  When the level has few nodes (the
// coarse levels) every pair lands on the same few mask words: a block-private
// copy in LDS takes the ORs and each block then adds only its non-zero words;
// otherwise the ORs go to global memory, combined per wave.  Grid-stride over
// the pairs; the path is uniform per block (it depends on the level's count).
constexpr int kMaskLdsWords = 8192;

template <int B>
__device__ __forceinline__ void contact_masks(int p, int np, int n, int coff, const int4* __restrict__ pairs,
                                              const int* __restrict__ map_l, int cv[4], uint32_t m[4])
{
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        cv[i] = -1;
        m[i] = 0;
    }
    if (p >= np) return;
    int v[4];
    decode_pair(pairs[p], coff, n, v);
#pragma unroll
    for (int i = 0; i < 4; ++i) cv[i] = v[i] >= 0 ? (map_l ? map_l[v[i]] : v[i]) : -1;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = i + 1; j < 4; ++j)
            if (cv[i] >= 0 && cv[j] >= 0 && cv[i] != cv[j] && cv[i] / B == cv[j] / B) {
                m[i] |= 1u << (cv[j] % B);
                m[j] |= 1u << (cv[i] % B);
            }
}

template <int B>
__global__ __launch_bounds__(256) void k_mask_contact(const int* __restrict__ meta, int level, int n, int coff,
                                                      const int4* __restrict__ pairs, int np,
                                                      const int* __restrict__ map_l, uint32_t* __restrict__ mask)
{
    __shared__ uint32_t sm[kMaskLdsWords];
    const int c = meta[kMetaCnt + level];
    const int stride = gridDim.x * blockDim.x;
    int cv[4];
    uint32_t m[4];
    if (c <= kMaskLdsWords) {
        for (int i = threadIdx.x; i < c; i += blockDim.x) sm[i] = 0u;
        __syncthreads();
        for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += stride) {
            contact_masks<B>(p, np, n, coff, pairs, map_l, cv, m);
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (cv[i] >= 0 && cv[i] < c && m[i]) atomicOr(sm + cv[i], m[i]);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < c; i += blockDim.x)
            if (sm[i]) atomicOr(mask + i, sm[i]);
    } else {
        // wave-uniform trip count: every lane takes part in the wave's combine
        for (int p0 = blockIdx.x * blockDim.x; p0 < np; p0 += stride) {
            contact_masks<B>(p0 + (int)threadIdx.x, np, n, coff, pairs, map_l, cv, m);
#pragma unroll
            for (int i = 0; i < 4; ++i) wave_or_atomic(mask, cv[i] >= 0 ? cv[i] : 0, cv[i] >= 0 ? m[i] : 0u);
        }
    }
}

// Connected components inside each domain (closure of the bitmasks) and the
// rank of every node's component among the domain's components.
template <int B>
__global__ __launch_bounds__(256) void k_closure(const int* __restrict__ meta, int level,
                                                 const uint32_t* __restrict__ mask,
                                                 int* __restrict__ comp_rank, int* __restrict__ dom_count)
{
    __shared__ uint32_t sm[256];
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = meta[kMetaCnt + level];
    const int lane = a % B;
    const bool real = a < c;
    uint32_t conn = real ? (mask[a] | (1u << lane)) : 0u;
    sm[threadIdx.x] = conn;
    __syncthreads();
    const int base = (threadIdx.x / B) * B;
    uint32_t visited = real ? (1u << lane) : 0u;
    while (true) {
        uint32_t todo = conn & ~visited;
        if (!todo) break;
        int nb = __ffs(todo) - 1;
        visited |= 1u << nb;
        conn |= sm[base + nb];
    }
    const int leader = real ? __ffs(conn) - 1 : -1;
    const bool is_leader = real && leader == lane;
    const uint64_t bal = __ballot(is_leader);
    const int shift = ((threadIdx.x & 63) / B) * B;
    const uint32_t dom_leaders =
        (uint32_t)((bal >> shift) & (B == 32 ? 0xffffffffull : 0xffffull));
    if (real) comp_rank[a] = __popc(dom_leaders & ((1u << leader) - 1u));
    const int ndom = (c + B - 1) / B;
    if (lane == 0 && a / B < ndom) dom_count[a / B] = __popc(dom_leaders);
}

// Exclusive scan of the per-domain component counts (single workgroup).
__global__ __launch_bounds__(1024) void k_scan_domains(int* __restrict__ meta, int level, int B,
                                                       const int* __restrict__ dom_count,
                                                       int* __restrict__ dom_prefix)
{
    __shared__ int ws[16];
    const int c = meta[kMetaCnt + level];
    const int ndom = (c + B - 1) / B;
    const int per = (ndom + blockDim.x - 1) / blockDim.x;
    const int beg = threadIdx.x * per;
    int end = beg + per;
    if (end > ndom) end = ndom;
    int local = 0;
    for (int d = beg; d < end; ++d) local += dom_count[d];
    // inclusive scan of `local` across the block
    int v = local;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    if (lane == 63) ws[w] = v;
    __syncthreads();
    if (w == 0) {
        int s = lane < (int)(blockDim.x >> 6) ? ws[lane] : 0;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            int t = __shfl_up(s, off, 64);
            if (lane >= off) s += t;
        }
        if (lane < 16) ws[lane] = s;
    }
    __syncthreads();
    int excl = v - local + (w > 0 ? ws[w - 1] : 0);
    for (int d = beg; d < end; ++d) {
        dom_prefix[d] = excl;
        excl += dom_count[d];
    }
    if (threadIdx.x == blockDim.x - 1) {
        int totalc = (w > 0 ? ws[w - 1] : 0) + v;
        meta[kMetaCnt + level + 1] = totalc;
        meta[kMetaOff + level + 1] = meta[kMetaOff + level] + ndom * B;
    }
}

__global__ void k_assign(const int* __restrict__ meta, int level, int B, const int* __restrict__ comp_rank,
                         const int* __restrict__ dom_prefix, int* __restrict__ next_id,
                         int* __restrict__ going_next)
{
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    int c = meta[kMetaCnt + level];
    if (a >= c) return;
    int id = dom_prefix[a / B] + comp_rank[a];
    next_id[a] = id;
    going_next[meta[kMetaOff + level] + a] = meta[kMetaOff + level + 1] + id;
}

__global__ void k_map_next(int n, const int* __restrict__ map_l, const int* __restrict__ next_id,
                           int* __restrict__ map_next)
{
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    int a = map_l ? map_l[u] : u;
    map_next[u] = next_id[a];
}

// row_ptr[i] = first block with local row >= i (rows sorted), i in [0, n].
__global__ void k_row_ptr(const int* __restrict__ rows, int nnz, int offset, int n, int* __restrict__ row_ptr)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    int key = i + offset, lo = 0, hi = nnz;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (rows[mid] < key) lo = mid + 1; else hi = mid;
    }
    row_ptr[i] = lo;
}

// rows and columns of the BCOO entries e0.. shifted down by offset
__global__ void k_shift_rc(const int* __restrict__ rows, const int* __restrict__ cols, long e0, int m, int offset,
                           int* __restrict__ sr, int* __restrict__ sc)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    sr[k] = rows[e0 + k] - offset;
    sc[k] = cols[e0 + k] - offset;
}

// ------------------------------------------------------------------ Galerkin
__device__ __forceinline__ void load_block(const double* __restrict__ v, double h[9])
{
#pragma unroll
    for (int e = 0; e < 9; ++e) h[e] = v[e];
}

// Galerkin coarsening of a full upper operator one level up:
// (i, j, H) with i <= j at level l-1 maps to (I, J) = (next(i), next(j)) at
// level l; H goes to (I, J), H^T to (J, I), and both halves land on the
// diagonal when I == J (H + H^T, or H alone for a diagonal block).  The
// mapped triplets are then merged by the deterministic converter, so
// A_l = P_l^T A_{l-1} P_l is formed without atomics.
__global__ __launch_bounds__(256) void k_coarsen_map(const double* __restrict__ vals, const int* __restrict__ rows,
                                                     const int* __restrict__ cols, int nnz, int offset, int n_prev,
                                                     const int* __restrict__ going_next, int off_prev, int off_next,
                                                     double* __restrict__ ov, int* __restrict__ orow,
                                                     int* __restrict__ ocol)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const int i = rows[k] - offset, j = cols[k] - offset;
    double* dst = ov + 9 * (long)k;
    if (i < 0 || j < 0 || i >= n_prev || j >= n_prev) {
        // outside the preconditioned range: contributes nothing
#pragma unroll
        for (int e = 0; e < 9; ++e) dst[e] = 0.0;
        orow[k] = 0;
        ocol[k] = 0;
        return;
    }
    const int I = going_next[off_prev + i] - off_next;
    const int J = going_next[off_prev + j] - off_next;
    double h[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) h[e] = vals[9 * (long)k + e];
    if (I == J) {
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) dst[c * 3 + r] = (i == j) ? h[c * 3 + r] : h[c * 3 + r] + h[r * 3 + c];
        orow[k] = I;
        ocol[k] = I;
    } else if (I < J) {
#pragma unroll
        for (int e = 0; e < 9; ++e) dst[e] = h[e];
        orow[k] = I;
        ocol[k] = J;
    } else {
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) dst[c * 3 + r] = h[r * 3 + c];
        orow[k] = J;
        ocol[k] = I;
    }
}

// Blocks of a level-l operator whose ends share a level-l domain become that
// domain's Galerkin blocks (assignment: the operator's blocks are unique).
template <int B>
__global__ __launch_bounds__(256) void k_extract_domains(const double* __restrict__ vals, const int* __restrict__ rows,
                                                         const int* __restrict__ cols, int nnz, int off_l, int ndom0,
                                                         double* __restrict__ mgal)
{
    constexpr int nblk = B * (B + 1) / 2;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const int I = rows[k], J = cols[k];
    if (I / B != J / B) return;
    const int dom = (off_l + I) / B - ndom0;
    double* dst = mgal + ((long)dom * nblk + blk_index(B, I % B, J % B)) * 9;
#pragma unroll
    for (int e = 0; e < 9; ++e) dst[e] = vals[9 * (long)k + e];
}

// ------------------------------------------------------------------ inverse
template <int B>
struct InvShape {
    static constexpr int N = 3 * B;
    static constexpr int G = N / 6;  // tile grid edge
    static constexpr int T = G * G;  // threads
};

// In-place Gauss-Jordan without pivoting on a 3Bx3B matrix distributed in 6x6
// register tiles (thread (ty,tx) holds rows 6ty.., cols 6tx..).  With
// nonpos != nullptr, a pivot that is not positive (the matrix is not positive
// definite) sets *nonpos.
template <int B>
__device__ __forceinline__ void gj_invert(double a[6][6], double* __restrict__ colbuf,
                                          double* __restrict__ rowbuf, int ty, int tx, int* nonpos = nullptr)
{
    constexpr int N = InvShape<B>::N;
    // zero diagonal -> 1 (padding slots), MASPreconditioner.cu:660-666
    if (ty == tx) {
#pragma unroll
        for (int d = 0; d < 6; ++d)
            if (a[d][d] == 0.0) a[d][d] = 1.0;
    }
    for (int j = 0; j < N; ++j) {
        double* cb = colbuf + (j & 1) * N;
        double* rb = rowbuf + (j & 1) * N;
        const int jt = j / 6, jo = j % 6;
        if (tx == jt) {
#pragma unroll
            for (int r = 0; r < 6; ++r)
                if (r == jo) {
#pragma unroll
                    for (int rr = 0; rr < 6; ++rr) cb[6 * ty + rr] = a[rr][r];
                }
        }
        if (ty == jt) {
#pragma unroll
            for (int r = 0; r < 6; ++r)
                if (r == jo) {
#pragma unroll
                    for (int cc = 0; cc < 6; ++cc) rb[6 * tx + cc] = a[r][cc];
                }
        }
        __syncthreads();
        const double ip = 1.0 / cb[j];
        if (nonpos != nullptr && threadIdx.x == 0 && !(cb[j] > 0.0)) *nonpos = 1;
        double rc[6], cr[6];
#pragma unroll
        for (int cc = 0; cc < 6; ++cc) rc[cc] = rb[6 * tx + cc] * ip;
#pragma unroll
        for (int rr = 0; rr < 6; ++rr) cr[rr] = cb[6 * ty + rr];
#pragma unroll
        for (int rr = 0; rr < 6; ++rr)
#pragma unroll
            for (int cc = 0; cc < 6; ++cc) a[rr][cc] -= cr[rr] * rc[cc];
        if (ty == jt) {  // pivot row: scaled
#pragma unroll
            for (int r = 0; r < 6; ++r)
                if (r == jo) {
#pragma unroll
                    for (int cc = 0; cc < 6; ++cc) a[r][cc] = rc[cc];
                }
        }
        if (tx == jt) {  // pivot column: -col/p, pivot 1/p
#pragma unroll
            for (int c = 0; c < 6; ++c)
                if (c == jo) {
#pragma unroll
                    for (int rr = 0; rr < 6; ++rr) a[rr][c] = -cr[rr] * ip;
                    if (ty == jt) {
#pragma unroll
                        for (int r = 0; r < 6; ++r)
                            if (r == jo) a[r][c] = ip;
                    }
                }
        }
    }
}

// Store the upper 3x3 blocks of the inverse (diagonal blocks symmetrised from
// their upper part, MASPreconditioner.cu:700-705).
template <int B, typename TO>
__device__ __forceinline__ void store_inverse(double a[6][6], int ty, int tx, TO* __restrict__ out)
{
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            const int rb = 2 * ty + bi, cb = 2 * tx + bj;
            if (rb > cb) continue;
            TO* dst = out + 9 * blk_index(B, rb, cb);
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    double v = a[3 * bi + r][3 * bj + c];
                    if (rb == cb && r > c) v = a[3 * bi + c][3 * bj + r];
                    dst[c * 3 + r] = (TO)v;
                }
        }
}

// Find block (u, v) (local ids, u <= v) in the sorted BCOO; false if absent.
__device__ __forceinline__ bool find_block(const int* __restrict__ row_ptr, const int* __restrict__ cols,
                                           int u, int v, int offset, long* pos)
{
    int lo = row_ptr[u], hi = row_ptr[u + 1];
    const int key = v + offset;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        int c = cols[mid];
        if (c == key) {
            *pos = mid;
            return true;
        }
        if (c < key) lo = mid + 1; else hi = mid;
    }
    return false;
}

// Level-0 domains: gather the domain's blocks straight from the sorted BCOO
// into register tiles, invert, store.
template <int B, typename TO>
__global__ __launch_bounds__(InvShape<B>::T) void k_inv_level0(int n, const double* __restrict__ vals,
                                                               const int* __restrict__ cols,
                                                               const int* __restrict__ row_ptr, int offset,
                                                               TO* __restrict__ minv)
{
    constexpr int N = InvShape<B>::N, G = InvShape<B>::G;
    constexpr int nblk = B * (B + 1) / 2;
    __shared__ double colbuf[2 * N], rowbuf[2 * N];
    const int d = blockIdx.x;
    const int ty = threadIdx.x / G, tx = threadIdx.x % G;
    double a[6][6];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            const int rb = 2 * ty + bi, cb = 2 * tx + bj;
            const int u = d * B + rb, v = d * B + cb;
            double h[9];
#pragma unroll
            for (int e = 0; e < 9; ++e) h[e] = 0.0;
            if (u < n && v < n) {
                long pos;
                if (rb <= cb) {
                    if (find_block(row_ptr, cols, u, v, offset, &pos)) load_block(vals + 9 * pos, h);
                } else {
                    if (find_block(row_ptr, cols, v, u, offset, &pos)) {
                        double t[9];
                        load_block(vals + 9 * pos, t);
#pragma unroll
                        for (int r = 0; r < 3; ++r)
#pragma unroll
                            for (int c = 0; c < 3; ++c) h[c * 3 + r] = t[r * 3 + c];
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) a[3 * bi + r][3 * bj + c] = h[c * 3 + r];
        }
    gj_invert<B>(a, colbuf, rowbuf, ty, tx);
    store_inverse<B, TO>(a, ty, tx, minv + (long)d * nblk * 9);
}

// Coarse domains: blocks from the Galerkin array.
template <int B, typename TO>
__global__ __launch_bounds__(InvShape<B>::T) void k_inv_coarse(const double* __restrict__ mgal, int ndom0,
                                                              TO* __restrict__ minv)
{
    constexpr int N = InvShape<B>::N, G = InvShape<B>::G;
    constexpr int nblk = B * (B + 1) / 2;
    __shared__ double colbuf[2 * N], rowbuf[2 * N];
    const int dc = blockIdx.x;
    const int ty = threadIdx.x / G, tx = threadIdx.x % G;
    const double* M = mgal + (long)dc * nblk * 9;
    double a[6][6];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            const int rb = 2 * ty + bi, cb = 2 * tx + bj;
            double h[9];
            if (rb <= cb) {
                load_block(M + 9 * blk_index(B, rb, cb), h);
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) a[3 * bi + r][3 * bj + c] = h[c * 3 + r];
            } else {
                load_block(M + 9 * blk_index(B, cb, rb), h);
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) a[3 * bi + r][3 * bj + c] = h[r * 3 + c];
            }
        }
    gj_invert<B>(a, colbuf, rowbuf, ty, tx);
    store_inverse<B, TO>(a, ty, tx, minv + (long)(ndom0 + dc) * nblk * 9);
}

// ------------------------------------------------------------------ apply
template <typename TS>
struct Vec3 {
    TS x, y, z;
};

// Level 0 -> slots and level 1 sums; evaluates the PCG stop rule first.
template <int B, typename TV, typename TS>
__global__ __launch_bounds__(256) void k_restrict0(PcgStopDesc sd, const int* __restrict__ meta, int n,
                                                   int L, const int* __restrict__ map1,
                                                   const TV* __restrict__ r, TS* __restrict__ R)
{
    __shared__ double smem[16];
    __shared__ int s_done;
    if (sd.ctl != nullptr) {
        // one read per workgroup: block 0 of this kernel may raise the flag
        if (threadIdx.x == 0) s_done = sd.ctl->done;
        __syncthreads();
        if (s_done) return;
        double rz = block_sum_array(sd.rz, sd.n_rz, smem);
        double rr = block_sum_array(sd.rr, sd.n_rr, smem);
        if (pcg_stop_verdict(sd.ctl, rz, rr)) return;
    }
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int npad = meta[kMetaOff + 1];
    const bool real = s < n;
    TS v0 = 0, v1 = 0, v2 = 0;
    if (real) {
        v0 = (TS)r[3 * (long)s];
        v1 = (TS)r[3 * (long)s + 1];
        v2 = (TS)r[3 * (long)s + 2];
    }
    if (s < npad) {
        R[3 * (long)s] = v0;
        R[3 * (long)s + 1] = v1;
        R[3 * (long)s + 2] = v2;
    }
    if (L > 1) {
        // group-by level-1 node inside the wave (a node lives in one domain)
        int node = real ? map1[s] : -1;
        bool pending = node >= 0;
        const int lane = threadIdx.x & 63;
        while (true) {
            const uint64_t act = __ballot(pending);
            if (act == 0) break;
            const int leader = __ffsll((unsigned long long)act) - 1;
            const int key = __shfl(node, leader, 64);
            const bool mine = pending && node == key;
            TS s0 = wave_sum(mine ? v0 : TS(0));
            TS s1 = wave_sum(mine ? v1 : TS(0));
            TS s2 = wave_sum(mine ? v2 : TS(0));
            if (lane == leader) {
                const long o = 3 * (long)(npad + key);
                R[o] = s0;
                R[o + 1] = s1;
                R[o + 2] = s2;
            }
            if (mine) pending = false;
        }
        // zero the padding slots of level 1
        const int c1 = meta[kMetaCnt + 1];
        const int pad = meta[kMetaOff + 2] - npad - c1;
        if (s < pad) {
            const long o = 3 * (long)(npad + c1 + s);
            R[o] = 0;
            R[o + 1] = 0;
            R[o + 2] = 0;
        }
    }
}

// x += alpha p, r -= alpha Ap (alpha = rz / pAp from partials), ||r||^2
// partials, then the restriction of the new r as in k_restrict0 (level 0
// only with SUMS = false: coarse mode 2 restricts r1 later).  The stop rule is
// left to the first coarse restriction, which sees the new rr.
template <int B, typename TV, typename TS, bool SUMS = true>
__global__ __launch_bounds__(256) void k_restrict0_xr(FusedXR<TV> fx, const int* __restrict__ meta, int n, int L,
                                                      const int* __restrict__ map1, TS* __restrict__ R)
{
    __shared__ double smem[16];
    __shared__ int s_done;
    if (threadIdx.x == 0) s_done = fx.ctl->done;
    __syncthreads();
    if (s_done) return;
    if constexpr (!SUMS) {
        // level 0 only: the update is elementwise over the 3 n entries, so a
        // block streams 768 consecutive entries (coalesced), its loads issued
        // before the partial sums that give alpha
        const long n3 = 3L * n;
        const long np3 = 3L * meta[kMetaOff + 1];
        const long base = (long)blockIdx.x * 768 + threadIdx.x;
        double rv[3], av[3], xv[3], pv[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const long k = base + 256 * j;
            const bool in = k < n3;
            rv[j] = in ? (double)fx.r[k] : 0.0;
            av[j] = in ? (double)fx.Ap[k] : 0.0;
            xv[j] = in ? (double)fx.x[k] : 0.0;
            pv[j] = in ? (double)fx.p[k] : 0.0;
        }
        const double rz = block_sum_array(fx.rz, fx.n_rz, smem);
        const double pap = block_sum_array(fx.pap, fx.n_pap, smem);
        const double alpha = rz / pap;
        double rrv = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const long k = base + 256 * j;
            if (k < n3) {
                const double ri = rv[j] - alpha * av[j];
                fx.x[k] = (TV)(xv[j] + alpha * pv[j]);
                fx.r[k] = (TV)ri;
                rrv += ri * ri;
                R[k] = (TS)(TV)ri;
            } else if (k < np3) {
                R[k] = 0;  // padding rows of level 0
            }
        }
        const double rr = block_sum(rrv, smem);
        if (threadIdx.x == 0) {
            fx.rr[blockIdx.x] = rr;
            if (blockIdx.x == 0) {
                fx.ctl->iters += 1;
                fx.ctl->last_alpha = alpha;
                if (!(pap > 0.0)) fx.ctl->not_spd = 1;
            }
        }
        return;
    }
    const double rz = block_sum_array(fx.rz, fx.n_rz, smem);
    const double pap = block_sum_array(fx.pap, fx.n_pap, smem);
    const double alpha = rz / pap;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int npad = meta[kMetaOff + 1];
    const bool real = s < n;
    TS v0 = 0, v1 = 0, v2 = 0;
    double rrv = 0.0;
    if (real) {
        double rv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const long k = 3 * (long)s + c;
            rv[c] = (double)fx.r[k] - alpha * (double)fx.Ap[k];
            fx.x[k] = (TV)((double)fx.x[k] + alpha * (double)fx.p[k]);
            fx.r[k] = (TV)rv[c];
            rrv += rv[c] * rv[c];
        }
        v0 = (TS)(TV)rv[0];
        v1 = (TS)(TV)rv[1];
        v2 = (TS)(TV)rv[2];
    }
    if (s < npad) {
        R[3 * (long)s] = v0;
        R[3 * (long)s + 1] = v1;
        R[3 * (long)s + 2] = v2;
    }
    if (SUMS) {
        int node = real ? map1[s] : -1;
        bool pending = node >= 0;
        const int lane = threadIdx.x & 63;
        while (true) {
            const uint64_t act = __ballot(pending);
            if (act == 0) break;
            const int leader = __ffsll((unsigned long long)act) - 1;
            const int key = __shfl(node, leader, 64);
            const bool mine = pending && node == key;
            TS s0 = wave_sum(mine ? v0 : TS(0));
            TS s1 = wave_sum(mine ? v1 : TS(0));
            TS s2 = wave_sum(mine ? v2 : TS(0));
            if (lane == leader) {
                const long o = 3 * (long)(npad + key);
                R[o] = s0;
                R[o + 1] = s1;
                R[o + 2] = s2;
            }
            if (mine) pending = false;
        }
        const int c1 = meta[kMetaCnt + 1];
        const int pad = meta[kMetaOff + 2] - npad - c1;
        if (s < pad) {
            const long o = 3 * (long)(npad + c1 + s);
            R[o] = 0;
            R[o + 1] = 0;
            R[o + 2] = 0;
        }
    }
    const double rr = block_sum(rrv, smem);
    if (threadIdx.x == 0) {
        fx.rr[blockIdx.x] = rr;
        if (blockIdx.x == 0) {
            fx.ctl->iters += 1;
            fx.ctl->last_alpha = alpha;
            if (!(pap > 0.0)) fx.ctl->not_spd = 1;
        }
    }
}

// Level l-1 -> level l sums (l >= 2).
template <typename TS>
__global__ __launch_bounds__(256) void k_restrict_up(const int* __restrict__ done, const int* __restrict__ meta,
                                                     int level, const int* __restrict__ going_next,
                                                     TS* __restrict__ R, PcgStopDesc sd)
{
    if (sd.ctl != nullptr) {
        // first coarse restriction after a fused x/r update: the stop rule
        __shared__ double smem[16];
        __shared__ int s_done;
        if (threadIdx.x == 0) s_done = sd.ctl->done;
        __syncthreads();
        if (s_done) return;
        const double rz = block_sum_array(sd.rz, sd.n_rz, smem);
        const double rr = block_sum_array(sd.rr, sd.n_rr, smem);
        if (pcg_stop_verdict(sd.ctl, rz, rr)) return;
    } else if (done != nullptr && *done) {
        return;
    }
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int cprev = meta[kMetaCnt + level - 1];
    const int offprev = meta[kMetaOff + level - 1];
    const bool real = a < cprev;
    TS v0 = 0, v1 = 0, v2 = 0;
    int parent = -1;
    if (real) {
        const long o = 3 * (long)(offprev + a);
        v0 = R[o];
        v1 = R[o + 1];
        v2 = R[o + 2];
        parent = going_next[offprev + a];
    }
    bool pending = real;
    const int lane = threadIdx.x & 63;
    while (true) {
        const uint64_t act = __ballot(pending);
        if (act == 0) break;
        const int leader = __ffsll((unsigned long long)act) - 1;
        const int key = __shfl(parent, leader, 64);
        const bool mine = pending && parent == key;
        TS s0 = wave_sum(mine ? v0 : TS(0));
        TS s1 = wave_sum(mine ? v1 : TS(0));
        TS s2 = wave_sum(mine ? v2 : TS(0));
        if (lane == leader) {
            R[3 * (long)key] = s0;
            R[3 * (long)key + 1] = s1;
            R[3 * (long)key + 2] = s2;
        }
        if (mine) pending = false;
    }
    const int c = meta[kMetaCnt + level];
    const int off = meta[kMetaOff + level];
    const int pad = meta[kMetaOff + level + 1] - off - c;
    if (a < pad) {
        const long o = 3 * (long)(off + c + a);
        R[o] = 0;
        R[o + 1] = 0;
        R[o + 2] = 0;
    }
}

template <typename TS>
struct Vec16;
template <>
struct Vec16<double> {
    using V = double2;
    static constexpr int k = 2;
};
template <>
struct Vec16<float> {
    using V = float4;
    static constexpr int k = 4;
};

template <typename TS>
struct Vec4Of;
template <>
struct Vec4Of<float> {
    using T = float4;
};
template <>
struct Vec4Of<double> {
    using T = double4;
};

// Row block rb of Minv r for a domain staged in LDS (upper 3x3 blocks, column
// major inside a block; r in sR): thread (rb, g) sums the column blocks
// cb = g, g+G, ..., and the G partials of a row, held by adjacent lanes, are
// added by shuffles in a fixed order.  Every lane ends with its row's sum.
template <int B, typename TS>
__device__ __forceinline__ void domain_product(const TS* __restrict__ sM, const TS* __restrict__ sR, int rb, int g,
                                               TS& z0, TS& z1, TS& z2)
{
    constexpr int G = 256 / B;
    z0 = 0;
    z1 = 0;
    z2 = 0;
#pragma unroll
    for (int cb = g; cb < B; cb += G) {
        const TS r0 = sR[3 * cb], r1 = sR[3 * cb + 1], r2 = sR[3 * cb + 2];
        if (cb >= rb) {
            const TS* m = sM + 9 * blk_index(B, rb, cb);
            z0 += m[0] * r0 + m[3] * r1 + m[6] * r2;
            z1 += m[1] * r0 + m[4] * r1 + m[7] * r2;
            z2 += m[2] * r0 + m[5] * r1 + m[8] * r2;
        } else {
            const TS* m = sM + 9 * blk_index(B, cb, rb);
            z0 += m[0] * r0 + m[1] * r1 + m[2] * r2;
            z1 += m[3] * r0 + m[4] * r1 + m[5] * r2;
            z2 += m[6] * r0 + m[7] * r1 + m[8] * r2;
        }
    }
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) {
        z0 += __shfl_xor(z0, off, 64);
        z1 += __shfl_xor(z1, off, 64);
        z2 += __shfl_xor(z2, off, 64);
    }
}

// FP16 storage of the level-0 inverse (coarse mode 2): C = D^-1 Minv D^-1
// with D = diag(Minv)^(1/2) per slot, so |C_ij| <= 1 for an SPD domain
// inverse; Minv r = D C D r.  16 bytes of storage (8 halves) unpack to two
// float4 units in LDS (order: half_slot_entry).
__device__ __forceinline__ float2 half2_bits_to_float2(unsigned u)
{
    __half2 h;
    __builtin_memcpy(&h, &u, sizeof(h));
    return __half22float2(h);
}

// Order of the halves in HBM.  A lane loads 16 bytes (8 halves, group u)
// and stores them to LDS as two float4 units; the 64 groups of a wave,
// c = u / 64, hold the float4 units 128 c + e (first four halves) and
// 128 c + m + e (last four), e = u % 64 and m the group count of that wave
// (64 but in the last), so each LDS store instruction covers consecutive
// units (no bank conflicts).  Returns the natural float index (block-major,
// 9 entries per block) of half h.
template <int B>
__host__ __device__ __forceinline__ int half_slot_entry(int h)
{
    constexpr int nval = B * (B + 1) / 2 * 9;
    constexpr int nvh = nval / 8;
    const int u = h / 8, j = h % 8;
    const int c = u / 64, e = u % 64;
    const int m = nvh - 64 * c < 64 ? nvh - 64 * c : 64;
    const int unit = 128 * c + (j < 4 ? e : m + e);
    return 4 * unit + (j & 3);
}

template <int B>
__global__ __launch_bounds__(256) void k_half_level0(const float* __restrict__ minv, uint16_t* __restrict__ minv16,
                                                      float* __restrict__ msc)
{
    constexpr int nblk = B * (B + 1) / 2;
    constexpr int nval = nblk * 9;
    __shared__ float sinv[3 * B];
    const int d = blockIdx.x;
    const float* src = minv + (long)d * nval;
    if (threadIdx.x < 3 * B) {
        const int rb = threadIdx.x / 3, i = threadIdx.x - 3 * (threadIdx.x / 3);
        const float v = src[9 * blk_index(B, rb, rb) + 4 * i];
        const bool pos = v > 0.0f;  // padding slots (and a broken domain) get scale 0
        const float sc = pos ? sqrtf(v) : 0.0f;
        msc[3L * d * B + threadIdx.x] = sc;
        sinv[threadIdx.x] = pos ? 1.0f / sc : 0.0f;
    }
    __syncthreads();
    __half* dst = reinterpret_cast<__half*>(minv16) + (long)d * nval;
    for (int h = threadIdx.x; h < nval; h += blockDim.x) {
        const int e = half_slot_entry<B>(h);
        const int blk = e / 9, k = e - 9 * (e / 9);
        int rb = 0;
        while (rb + 1 < B && blk_index(B, rb + 1, rb + 1) <= blk) ++rb;
        const int cb = blk - (B * rb - rb * (rb + 1) / 2);
        // block entry k is M(3 rb + k % 3, 3 cb + k / 3)
        float c = src[e] * sinv[3 * rb + k % 3] * sinv[3 * cb + k / 3];
        c = fminf(fmaxf(c, -65504.0f), 65504.0f);
        dst[h] = __float2half_rn(c);
    }
}

// Z[domain] = Minv[domain] * R[domain], one 256-thread workgroup per domain.
// All loads of the domain's inverse blocks are issued before any LDS store so
// the whole 19/38 KB domain is in flight at once.  Thread (rb, g) forms the
// partial product of block row rb with the column blocks cb = g, g+G, ...;
// the G partials of a row sit in adjacent lanes and are summed by shuffles in
// a fixed order.
template <int B, typename TS>
__global__ __launch_bounds__(256) void k_local_solve(const int* __restrict__ done, const TS* __restrict__ minv,
                                                      const TS* __restrict__ R, TS* __restrict__ Z, int dom0,
                                                      const TS* __restrict__ C2 = nullptr, void* __restrict__ c4 = nullptr,
                                                      int off1 = 0, int c1 = 0)
{
    if (done != nullptr && *done) return;
    constexpr int nblk = B * (B + 1) / 2;
    constexpr int nval = nblk * 9;
    using V = typename Vec16<TS>::V;
    constexpr int kv = Vec16<TS>::k;
    constexpr int nvec = nval / kv;
    constexpr int per = (nvec + 255) / 256;
    constexpr int G = 256 / B;  // threads per block row
    __shared__ __align__(16) TS sM[nval];
    __shared__ TS sR[3 * B];
    const int d = dom0 + blockIdx.x;
    const V* vsrc = reinterpret_cast<const V*>(minv + (long)d * nval);
    V tmp[per];
    // unconditional (clamped) loads so all of them are in flight together
#pragma unroll
    for (int q = 0; q < per; ++q) {
        const int e = threadIdx.x + 256 * q;
        tmp[q] = vsrc[e < nvec ? e : nvec - 1];
    }
    __builtin_amdgcn_sched_barrier(0);  // keep every load ahead of the LDS stores
    if (threadIdx.x < 3 * B) sR[threadIdx.x] = R[3 * (long)d * B + threadIdx.x];
#pragma unroll
    for (int q = 0; q < per; ++q) {
        const int e = threadIdx.x + 256 * q;
        if (e < nvec) reinterpret_cast<V*>(sM)[e] = tmp[q];
    }
    __syncthreads();
    const int rb = threadIdx.x / G, g = threadIdx.x % G;
    TS z0, z1, z2;
    domain_product<B>(sM, sR, rb, g, z0, z1, z2);
    if (g == 0) {
        const long o = 3 * ((long)d * B + rb);
        Z[o] = z0;
        Z[o + 1] = z1;
        Z[o + 2] = z2;
        // level-1 post-smoothing of coarse mode 2: the cycle's correction
        // C2 + Z1, packed per level-1 node for the post-smoothing's gathers
        if (c4 != nullptr) {
            const long k = (long)d * B + rb - off1;
            if (k >= 0 && k < c1) {
                typename Vec4Of<TS>::T v;
                v.x = C2[3 * k] + z0;
                v.y = C2[3 * k + 1] + z1;
                v.z = C2[3 * k + 2] + z2;
                v.w = 0;
                static_cast<typename Vec4Of<TS>::T*>(c4)[k] = v;
            }
        }
    }
}

constexpr int kPostChunk = 256;  // A P entries per pass of post_rhs

// Right-hand side of the level-0 post-smoothing for slot component
// threadIdx.x < 3B of domain d: r_u + R0_u - (A P c)_u (R0 holds r1, c4 the
// level-1 correction).  The domain's A P entries are contiguous (CSR rows in
// order) and load coalesced; each thread forms one entry's product into sP
// per pass, and each row component sums its row's entries in CSR order.
// Contains barriers: every thread of the workgroup calls it.
template <int B, typename TV, typename TS>
__device__ __forceinline__ TS post_rhs(int d, int n, const int* __restrict__ ap_ptr, const int* __restrict__ ap_col,
                                       const float* __restrict__ ap_blk, const typename Vec4Of<TS>::T* __restrict__ c4,
                                       const TV* __restrict__ r, const TS* __restrict__ R, TS* __restrict__ sP)
{
    const int u0 = d * B;
    const int u1 = u0 + B < n ? u0 + B : n;
    const int e0 = u0 < n ? ap_ptr[u0] : 0, e1 = u0 < n ? ap_ptr[u1] : 0;
    const int my_row = threadIdx.x / 3, my_c = threadIdx.x - 3 * (threadIdx.x / 3);
    const int u = u0 + my_row;
    const bool own = threadIdx.x < 3 * B && u < n;
    const int rb0 = own ? ap_ptr[u] : 0, rb1 = own ? ap_ptr[u + 1] : 0;
    TS acc = 0;
    for (int c0 = e0; c0 < e1; c0 += kPostChunk) {
        const int e = c0 + (int)threadIdx.x;
        if (e < e1) {
            const long g = ap_col[e];
            const float* b = ap_blk + 9L * e;
            const typename Vec4Of<TS>::T cv = c4[g];
            sP[3 * threadIdx.x] = (TS)b[0] * cv.x + (TS)b[3] * cv.y + (TS)b[6] * cv.z;
            sP[3 * threadIdx.x + 1] = (TS)b[1] * cv.x + (TS)b[4] * cv.y + (TS)b[7] * cv.z;
            sP[3 * threadIdx.x + 2] = (TS)b[2] * cv.x + (TS)b[5] * cv.y + (TS)b[8] * cv.z;
        }
        __syncthreads();
        if (own) {
            const int k0 = rb0 > c0 ? rb0 : c0, k1 = rb1 < c0 + kPostChunk ? rb1 : c0 + kPostChunk;
            for (int k = k0; k < k1; ++k) acc += sP[3 * (k - c0) + my_c];
        }
        __syncthreads();
    }
    if (!own) return TS(0);
    const long o = 3L * u0 + threadIdx.x;
    return (TS)r[o] + R[o] - acc;
}

// The right-hand side of the level-0 post-smoothing (r + r1 - A P c, zero on
// padding rows) written over R's level-0 slots, for the FP64 inverse, whose
// post-smoothing then runs as a plain k_local_solve: in one kernel the A P
// partials' LDS costs the FP64 inverse a workgroup per CU (3 instead of 4),
// more than the right-hand side's round trip through HBM.  Each thread reads
// and writes only its own slot component.
template <int B, typename TV, typename TS>
__global__ __launch_bounds__(256) void k_post_rhs(const int* __restrict__ done, TS* __restrict__ R, int n,
                                                  const int* __restrict__ ap_ptr, const int* __restrict__ ap_col,
                                                  const float* __restrict__ ap_blk,
                                                  const typename Vec4Of<TS>::T* __restrict__ c4,
                                                  const TV* __restrict__ r)
{
    if (done != nullptr && *done) return;
    __shared__ TS sP[3 * kPostChunk];
    const int d = blockIdx.x;
    const TS rhs = post_rhs<B>(d, n, ap_ptr, ap_col, ap_blk, c4, r, R, sP);
    if (threadIdx.x < 3 * B) R[3L * d * B + threadIdx.x] = rhs;
}

// Level-0 post-smoothing of coarse mode 2 with its right-hand side formed in
// the kernel (post_rhs: R2 = r + r1 - A P c), then Z = Minv R2 as in
// k_local_solve.  The inverse blocks are loaded after the A P phase, so no
// register holds them across its barriers.
template <int B, typename TV, typename TS>
__global__ __launch_bounds__(256) void k_local_solve_post(const int* __restrict__ done, const TS* __restrict__ minv,
                                                           const TS* __restrict__ R, TS* __restrict__ Z, int n,
                                                           const int* __restrict__ ap_ptr,
                                                           const int* __restrict__ ap_col,
                                                           const float* __restrict__ ap_blk,
                                                           const typename Vec4Of<TS>::T* __restrict__ c4,
                                                           const TV* __restrict__ r)
{
    if (done != nullptr && *done) return;
    constexpr int nblk = B * (B + 1) / 2;
    constexpr int nval = nblk * 9;
    using V = typename Vec16<TS>::V;
    constexpr int kv = Vec16<TS>::k;
    constexpr int nvec = nval / kv;
    constexpr int per = (nvec + 255) / 256;
    constexpr int G = 256 / B;
    __shared__ __align__(16) TS sM[nval];
    __shared__ TS sR[3 * B];
    __shared__ TS sP[3 * kPostChunk];
    const int d = blockIdx.x;
    const TS rhs = post_rhs<B>(d, n, ap_ptr, ap_col, ap_blk, c4, r, R, sP);
    if (threadIdx.x < 3 * B) sR[threadIdx.x] = rhs;
    // the domain's inverse blocks (loaded after the A P phase, so nothing is
    // held in registers across its barriers)
    const V* vsrc = reinterpret_cast<const V*>(minv + (long)d * nval);
    V tmp[per];
#pragma unroll
    for (int q = 0; q < per; ++q) {
        const int e = threadIdx.x + 256 * q;
        tmp[q] = vsrc[e < nvec ? e : nvec - 1];
    }
#pragma unroll
    for (int q = 0; q < per; ++q) {
        const int e = threadIdx.x + 256 * q;
        if (e < nvec) reinterpret_cast<V*>(sM)[e] = tmp[q];
    }
    __syncthreads();
    const int rb = threadIdx.x / G, g = threadIdx.x % G;
    TS z0, z1, z2;
    domain_product<B>(sM, sR, rb, g, z0, z1, z2);
    if (g == 0) {
        const long o = 3 * ((long)d * B + rb);
        Z[o] = z0;
        Z[o + 1] = z1;
        Z[o + 2] = z2;
    }
}

// The two level-0 solves of coarse mode 2 over the FP16 form of the
// inverse (k_half_level0): Minv R = D C D R.  A workgroup walks the domains
// d = blockIdx.x, + gridDim.x, ... (one domain each when the grid covers
// them); while one domain's product runs, the next domain's 9.5 KB of halves
// are already in flight.
template <int B>
struct HalfDomain {
    static constexpr int nval = B * (B + 1) / 2 * 9;
    static constexpr int nvh = nval / 8;  // 16-byte groups of 8 halves
    static constexpr int per = (nvh + 255) / 256;
    uint4 h[per];
    __device__ __forceinline__ void fetch(const uint4* __restrict__ minv16, int d)
    {
        const uint4* src = minv16 + (long)d * nvh;
#pragma unroll
        for (int q = 0; q < per; ++q) {
            const int e = threadIdx.x + 256 * q;
            h[q] = src[e < nvh ? e : nvh - 1];
        }
    }
    // the two float4 units of group u (the order of half_slot_entry)
    __device__ __forceinline__ void stage(float* __restrict__ sM) const
    {
#pragma unroll
        for (int q = 0; q < per; ++q) {
            const int u = threadIdx.x + 256 * q;
            if (u < nvh) {
                const int c = u / 64, e = u % 64;
                const int m = nvh - 64 * c < 64 ? nvh - 64 * c : 64;
                float4* dst = reinterpret_cast<float4*>(sM);
                const float2 a = half2_bits_to_float2(h[q].x), b = half2_bits_to_float2(h[q].y);
                const float2 cc = half2_bits_to_float2(h[q].z), dd = half2_bits_to_float2(h[q].w);
                dst[128 * c + e] = make_float4(a.x, a.y, b.x, b.y);
                dst[128 * c + m + e] = make_float4(cc.x, cc.y, dd.x, dd.y);
            }
        }
    }
};

template <int B>
__global__ __launch_bounds__(256) void k_solve0_half(const int* __restrict__ done, const uint4* __restrict__ minv16,
                                                      const float* __restrict__ msc, const float* __restrict__ R,
                                                      float* __restrict__ Z, int ndom)
{
    if (done != nullptr && *done) return;
    constexpr int G = 256 / B;
    __shared__ __align__(16) float sM[HalfDomain<B>::nval];
    __shared__ float sR[3 * B];
    const int t = threadIdx.x;
    const int rb = t / G, g = t % G;
    HalfDomain<B> hd;
    float rv = 0.0f;
    int d = blockIdx.x;
    if (d < ndom) {
        hd.fetch(minv16, d);
        if (t < 3 * B) rv = R[3L * d * B + t] * msc[3L * d * B + t];
    }
    for (; d < ndom; d += gridDim.x) {
        if (t < 3 * B) sR[t] = rv;
        hd.stage(sM);
        __syncthreads();
        const int dn = d + (int)gridDim.x;
        if (dn < ndom) {
            hd.fetch(minv16, dn);
            if (t < 3 * B) rv = R[3L * dn * B + t] * msc[3L * dn * B + t];
        }
        // rb and g made opaque per domain: the product's LDS offsets are
        // recomputed instead of held in registers across the loop
        int rbi = rb, gi = g;
        asm volatile("" : "+v"(rbi), "+v"(gi));
        float z0, z1, z2;
        domain_product<B>(sM, sR, rbi, gi, z0, z1, z2);
        if (gi == 0) {
            const long o = 3 * ((long)d * B + rbi);
            Z[o] = z0 * msc[o];
            Z[o + 1] = z1 * msc[o + 1];
            Z[o + 2] = z2 * msc[o + 2];
        }
        __syncthreads();  // sM and sR are restaged for the next domain
    }
}

template <int B, typename TV>
__global__ __launch_bounds__(256) void k_solve0_post_half(const int* __restrict__ done,
                                                           const uint4* __restrict__ minv16,
                                                           const float* __restrict__ msc, const float* __restrict__ R,
                                                           float* __restrict__ Z, int n, int ndom,
                                                           const int* __restrict__ ap_ptr,
                                                           const int* __restrict__ ap_col,
                                                           const float* __restrict__ ap_blk,
                                                           const float4* __restrict__ c4, const TV* __restrict__ r)
{
    if (done != nullptr && *done) return;
    constexpr int G = 256 / B;
    __shared__ __align__(16) float sM[HalfDomain<B>::nval];
    __shared__ float sR[3 * B];
    __shared__ float sP[3 * kPostChunk];
    const int t = threadIdx.x;
    const int rb = t / G, g = t % G;
    HalfDomain<B> hd;
    int d = blockIdx.x;
    if (d < ndom) hd.fetch(minv16, d);  // in flight during the A P phase
    for (; d < ndom; d += gridDim.x) {
        const float rhs = post_rhs<B>(d, n, ap_ptr, ap_col, ap_blk, c4, r, R, sP);
        if (t < 3 * B) sR[t] = rhs * msc[3L * d * B + t];
        hd.stage(sM);
        __syncthreads();
        const int dn = d + (int)gridDim.x;
        if (dn < ndom) hd.fetch(minv16, dn);
        // rb and g made opaque per domain: the product's LDS offsets are
        // recomputed instead of held in registers across the loop
        int rbi = rb, gi = g;
        asm volatile("" : "+v"(rbi), "+v"(gi));
        float z0, z1, z2;
        domain_product<B>(sM, sR, rbi, gi, z0, z1, z2);
        if (gi == 0) {
            const long o = 3 * ((long)d * B + rbi);
            Z[o] = z0 * msc[o];
            Z[o + 1] = z1 * msc[o + 1];
            Z[o + 2] = z2 * msc[o + 2];
        }
        __syncthreads();
    }
}

// z[v] = w0 Z[v] + sum_{l>=1} Z[slot_l(v)], with r.z partials (w0 = 1 but
// for the two-grid cycle of coarse mode 2, where it is the smoothing weight).  With a level-1 cycle
// (C2 != nullptr) the coarse part is C2[map1(v)] + Z[slot_1(v)] (the cycle's
// coarse-corrected iterate and its post-smoothing).  corr (level-1, float)
// adds the learned correction; with Ap set the z.Ap partials of flexible PCG
// go to zap_part.
template <typename TV, typename TS>
__global__ __launch_bounds__(256) void k_prolong(const int* __restrict__ done, const int* __restrict__ meta,
                                                 int n, int L, const int* __restrict__ map,
                                                 const TS* __restrict__ Z, const TV* __restrict__ r,
                                                 TV* __restrict__ z, double* __restrict__ rz_part,
                                                 const TS* __restrict__ C2, const float* __restrict__ corr,
                                                 const TV* __restrict__ Ap, double* __restrict__ zap_part, TS w0)
{
    __shared__ double smem[16];
    if (done != nullptr && *done) return;
    int offs[kMasMaxLevels];
#pragma unroll
    for (int l = 0; l < kMasMaxLevels; ++l) offs[l] = l < L ? meta[kMetaOff + l] : 0;
    double acc = 0.0, acc2 = 0.0;
    for (long u = (long)blockIdx.x * blockDim.x + threadIdx.x; u < n; u += (long)gridDim.x * blockDim.x) {
        TS z0 = w0 * Z[3 * u], z1 = w0 * Z[3 * u + 1], z2 = w0 * Z[3 * u + 2];
        if (C2 != nullptr) {
            const long k = map[(long)n + u];
            const long o = 3 * (offs[1] + k);
            z0 += C2[3 * k] + Z[o];
            z1 += C2[3 * k + 1] + Z[o + 1];
            z2 += C2[3 * k + 2] + Z[o + 2];
        } else {
            for (int l = 1; l < L; ++l) {
                const long o = 3 * (long)(offs[l] + map[(long)l * n + u]);
                z0 += Z[o];
                z1 += Z[o + 1];
                z2 += Z[o + 2];
            }
        }
        if (corr != nullptr) {
            const long k = map[(long)n + u];
            z0 += (TS)corr[3 * k];
            z1 += (TS)corr[3 * k + 1];
            z2 += (TS)corr[3 * k + 2];
        }
        const TV w0 = (TV)z0, w1 = (TV)z1, w2 = (TV)z2;
        z[3 * u] = w0;
        z[3 * u + 1] = w1;
        z[3 * u + 2] = w2;
        acc += (double)r[3 * u] * (double)w0 + (double)r[3 * u + 1] * (double)w1 + (double)r[3 * u + 2] * (double)w2;
        if (Ap != nullptr)
            acc2 += (double)Ap[3 * u] * (double)w0 + (double)Ap[3 * u + 1] * (double)w1 +
                    (double)Ap[3 * u + 2] * (double)w2;
    }
    double s = block_sum(acc, smem);
    if (threadIdx.x == 0 && rz_part) rz_part[blockIdx.x] = s;
    if (Ap != nullptr) {
        s = block_sum(acc2, smem);
        if (threadIdx.x == 0) zap_part[blockIdx.x] = s;
    }
}

// Level-1 residual of the coarse cycle, rho = g - A_1 c with
// A_1 = S^-1 Ahat S^-1 (the Jacobi-scaled operator of the level-1 graph),
// written over the level-1 slots of R.  One wave per level-1 node, lanes over
// its CSR row.  First pass: c = Z at the level-1 slots (the pre-smoothed
// iterate), g read from R and saved in G1.  Second pass: c = C2, g from G1.
// CHAIN (second pass): C2 is not read but formed here for every neighbour j
// -- the pre-smoothed level-1 iterate plus the coarse correction of levels
// 2 .. Lc-1 along going_next, added in that order (l1_chained) -- and
// written for the wave's own node.
template <typename TS>
__device__ __forceinline__ void l1_chained(const TS* __restrict__ Z, const int* __restrict__ going_next, int Lc,
                                           long sl, TS& a, TS& b, TS& c)
{
    a = Z[3 * sl];
    b = Z[3 * sl + 1];
    c = Z[3 * sl + 2];
    for (int l = 2; l < Lc; ++l) {
        sl = going_next[sl];
        a += Z[3 * sl];
        b += Z[3 * sl + 1];
        c += Z[3 * sl + 2];
    }
}

template <typename TS, bool SECOND, bool CHAIN = false>
__global__ __launch_bounds__(256) void k_l1_residual(const int* __restrict__ done, const int* __restrict__ meta, int c1,
                                                     const int* __restrict__ rp, const int* __restrict__ colx,
                                                     const float* __restrict__ blk, const float* __restrict__ sc,
                                                     TS* __restrict__ R, const TS* __restrict__ Z,
                                                     TS* __restrict__ G1, TS* __restrict__ C2,
                                                     const int* __restrict__ going_next = nullptr, int Lc = 0)
{
    static_assert(!CHAIN || SECOND, "the chain is formed in the second pass");
    if (done != nullptr && *done) return;
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= c1) return;
    const int off1 = meta[kMetaOff + 1];
    const TS* cv = SECOND ? C2 : Z + 3L * off1;
    TS a0 = 0, a1 = 0, a2 = 0;
    for (int e = rp[k] + lane; e < rp[k + 1]; e += 64) {
        const int j = colx[e];
        const float* b = blk + 9L * e;
        const float sj = sc[j];
        const TS wj = sj > 0.f ? (TS)1 / (TS)sj : (TS)0;
        TS x0, x1, x2;
        if constexpr (CHAIN) {
            l1_chained(Z, going_next, Lc, (long)off1 + j, x0, x1, x2);
        } else {
            x0 = cv[3L * j];
            x1 = cv[3L * j + 1];
            x2 = cv[3L * j + 2];
        }
        const TS c0 = wj * x0, c1v = wj * x1, c2v = wj * x2;
        a0 += (TS)b[0] * c0 + (TS)b[1] * c1v + (TS)b[2] * c2v;
        a1 += (TS)b[3] * c0 + (TS)b[4] * c1v + (TS)b[5] * c2v;
        a2 += (TS)b[6] * c0 + (TS)b[7] * c1v + (TS)b[8] * c2v;
    }
    a0 = wave_sum(a0);
    a1 = wave_sum(a1);
    a2 = wave_sum(a2);
    if (lane != 0) return;
    if constexpr (CHAIN) {
        TS x0, x1, x2;
        l1_chained(Z, going_next, Lc, (long)off1 + k, x0, x1, x2);
        C2[3L * k] = x0;
        C2[3L * k + 1] = x1;
        C2[3L * k + 2] = x2;
    }
    const long o = 3 * (long)(off1 + k);
    TS g0, g1, g2;
    if (!SECOND) {
        g0 = R[o];
        g1 = R[o + 1];
        g2 = R[o + 2];
        G1[3L * k] = g0;
        G1[3L * k + 1] = g1;
        G1[3L * k + 2] = g2;
    } else {
        g0 = G1[3L * k];
        g1 = G1[3L * k + 1];
        g2 = G1[3L * k + 2];
    }
    const float sk = sc[k];
    const TS wk = sk > 0.f ? (TS)1 / (TS)sk : (TS)0;
    R[o] = g0 - wk * a0;
    R[o + 1] = g1 - wk * a1;
    R[o + 2] = g2 - wk * a2;
}

// Inputs of the learned correction per level-1 node k (one wave per node):
// ghat = s_k g_k, zhat = zc_k / s_k and rho = ghat - Ahat zhat, where g is
// the level-1 residual sum and zc the coarse part of the MAS output at level
// 1 (levels >= 1 along going_next, or C2 + Z_1 with the level-1 cycle).
template <typename TS>
__global__ __launch_bounds__(256) void k_l1_features(const int* __restrict__ done, const int* __restrict__ meta, int L,
                                                     const int* __restrict__ going_next, int c1,
                                                     const int* __restrict__ rp, const int* __restrict__ colx,
                                                     const float* __restrict__ blk, const float* __restrict__ sc,
                                                     const TS* __restrict__ R, const TS* __restrict__ Z,
                                                     const TS* __restrict__ G1, const TS* __restrict__ C2,
                                                     float* __restrict__ X)
{
    if (done != nullptr && *done) return;
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= c1) return;
    const int off1 = meta[kMetaOff + 1];
    auto zc = [&](int j, double& a, double& b, double& c) {
        long sl = off1 + j;
        a = (double)Z[3 * sl];
        b = (double)Z[3 * sl + 1];
        c = (double)Z[3 * sl + 2];
        if (C2 != nullptr) {
            a += (double)C2[3L * j];
            b += (double)C2[3L * j + 1];
            c += (double)C2[3L * j + 2];
        } else {
            for (int l = 2; l < L; ++l) {
                sl = going_next[sl];
                a += (double)Z[3 * sl];
                b += (double)Z[3 * sl + 1];
                c += (double)Z[3 * sl + 2];
            }
        }
    };
    double a0 = 0, a1 = 0, a2 = 0;
    for (int e = rp[k] + lane; e < rp[k + 1]; e += 64) {
        const int j = colx[e];
        const float* b = blk + 9L * e;
        const float sj = sc[j];
        const double wj = sj > 0.f ? 1.0 / (double)sj : 0.0;
        double c0, c1v, c2v;
        zc(j, c0, c1v, c2v);
        c0 *= wj;
        c1v *= wj;
        c2v *= wj;
        a0 += (double)b[0] * c0 + (double)b[1] * c1v + (double)b[2] * c2v;
        a1 += (double)b[3] * c0 + (double)b[4] * c1v + (double)b[5] * c2v;
        a2 += (double)b[6] * c0 + (double)b[7] * c1v + (double)b[8] * c2v;
    }
    a0 = wave_sum(a0);
    a1 = wave_sum(a1);
    a2 = wave_sum(a2);
    if (lane != 0) return;
    const TS* gp = G1 != nullptr ? G1 + 3L * k : R + 3L * (off1 + k);
    const double sk = (double)sc[k];
    const double wk = sk > 0.0 ? 1.0 / sk : 0.0;
    double z0, z1, z2;
    zc(k, z0, z1, z2);
    const double g[3] = {sk * (double)gp[0], sk * (double)gp[1], sk * (double)gp[2]};
    const double zh[3] = {wk * z0, wk * z1, wk * z2};
    const double ah[3] = {a0, a1, a2};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float* x = X + (3L * k + d) * 3;
        x[0] = (float)g[d];
        x[1] = (float)zh[d];
        x[2] = (float)(g[d] - ah[d]);
    }
}

// Per-sample normalisation of training features: divide X (and write q =
// the rho column) by the RMS of the ghat column; one workgroup, fixed-order
// sums.
__global__ __launch_bounds__(1024) void k_sample_normalise(int rows, float* __restrict__ X, float* __restrict__ q)
{
    __shared__ double sm[16];
    double acc = 0.0;
    for (int i = threadIdx.x; i < rows; i += blockDim.x) acc += (double)X[3L * i] * (double)X[3L * i];
    const double ss = block_sum(acc, sm);
    const double rms = sqrt(ss / (double)(rows > 0 ? rows : 1));
    const float inv = rms > 0.0 ? (float)(1.0 / rms) : 1.f;
    for (int i = threadIdx.x; i < rows; i += blockDim.x) {
        X[3L * i] *= inv;
        X[3L * i + 1] *= inv;
        X[3L * i + 2] *= inv;
        q[i] = X[3L * i + 2];
    }
}

// ---- coarse mode 2: A P of the post-smoothing residual ----------------------
// (row u, level-1 node g) keys of A P: each upper block (i, j) adds A_ij to
// (i, map1[j]) and, off the diagonal, A_ij^T to (j, map1[i]).  Entry e = 2b + t
// (t = 1: transposed); the unused transposed entry of a diagonal block gets
// row n, which sorts after every real key.
// Keys (row << cb) | level-1 column, cb the bits of the level-1 count, so the
// sort covers only the bits in use.
__global__ void k_ap_keys(const int* __restrict__ rows, const int* __restrict__ cols, int nnz, int n,
                          const int* __restrict__ map1, int cb, uint64_t* __restrict__ keys, int* __restrict__ idx)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nnz) return;
    const int i = rows[b], j = cols[b];
    keys[2L * b] = ((uint64_t)(uint32_t)i << cb) | (uint32_t)map1[j];
    keys[2L * b + 1] = i != j ? (((uint64_t)(uint32_t)j << cb) | (uint32_t)map1[i]) : ((uint64_t)(uint32_t)n << cb);
    idx[2L * b] = 2 * b;
    idx[2L * b + 1] = 2 * b + 1;
}

// first entry of every distinct real key
__global__ void k_ap_flags(const uint64_t* __restrict__ k, int m, int n, int cb, int* __restrict__ flag)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    flag[e] = ((long)(k[e] >> cb) < n && (e == 0 || k[e] != k[e - 1])) ? 1 : 0;
}

__global__ void k_ap_cols(const uint64_t* __restrict__ k, const int* __restrict__ flag, const int* __restrict__ inc,
                          int m, int cb, int* __restrict__ col, int* __restrict__ seg_start)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m || !flag[e]) return;
    col[inc[e] - 1] = (int)(k[e] & ((1ull << cb) - 1ull));
    seg_start[inc[e] - 1] = e;
}

// ap_ptr[u] = first entry of row u (lower bound of u << cb among the sorted
// keys); for u = n the lower bound is the count of real keys, which ends the
// last segment (seg_start[n_uniq])
__global__ void k_ap_ptr(const uint64_t* __restrict__ k, const int* __restrict__ inc, int m, int n, int cb,
                         int* __restrict__ ptr, int* __restrict__ seg_start)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u > n) return;
    const int n_uniq = inc[m - 1];
    const uint64_t key = (uint64_t)(uint32_t)u << cb;
    int lo = 0, hi = m;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (k[mid] < key) lo = mid + 1; else hi = mid;
    }
    ptr[u] = (lo < m && (long)(k[lo] >> cb) < n) ? inc[lo] - 1 : n_uniq;
    if (u == n) seg_start[n_uniq] = lo;
}

// every distinct (u, g): its blocks summed in sorted order in FP64, stored in
// FP32 (column-major 3x3, as the BCOO)
__global__ void k_ap_sum(const double* __restrict__ vals, const int* __restrict__ idx, const int* __restrict__ seg_start,
                         const int* __restrict__ inc, int m, float* __restrict__ blk)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= inc[m - 1]) return;  // the distinct (row, level-1 column) pairs
    const int e0 = seg_start[q], e1 = seg_start[q + 1];
    double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int e = e0; e < e1; ++e) {
        const int v = idx[e];
        const double* h = vals + 9L * (v >> 1);
        if (v & 1) {
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) a[c * 3 + r] += h[r * 3 + c];
        } else {
#pragma unroll
            for (int t = 0; t < 9; ++t) a[t] += h[t];
        }
    }
#pragma unroll
    for (int t = 0; t < 9; ++t) blk[9L * q + t] = (float)a[t];
}

// y = Minv x for the dense FP64 inverse of the level-2 operator (m x m of a
// row-major array with row stride ld, symmetric); one workgroup per row (the
// matrix is small: many short rows in flight at once hide the load latency),
// summed in a fixed order.
template <typename TS>
__global__ __launch_bounds__(256) void k_dense_gemv(const int* __restrict__ done, int m, int ld,
                                                    const double* __restrict__ Minv, const TS* __restrict__ x,
                                                    TS* __restrict__ y)
{
    __shared__ double smem[16];
    if (done != nullptr && *done) return;
    const int row = blockIdx.x;
    const double* a = Minv + (long)row * ld;
    double acc = 0.0;
    for (int c = threadIdx.x; c < m; c += blockDim.x) acc += a[c] * (double)x[c];
    acc = block_sum(acc, smem);
    if (threadIdx.x == 0) y[row] = (TS)acc;
}

// Dense symmetric level-2 operator from its upper BCOO (column-major blocks)
// into rows/cols [0, m) of the ld x ld array D (zero-filled beforehand).
__global__ void k_dense_scatter(const double* __restrict__ v, const int* __restrict__ r, const int* __restrict__ c,
                                int nb, int ld, double* __restrict__ D)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    const long i0 = 3L * r[k], j0 = 3L * c[k];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            const double x = v[9L * k + 3 * b + a];  // M[a][b]
            D[(i0 + a) * ld + j0 + b] = x;
            D[(j0 + b) * ld + i0 + a] = x;
        }
}

// Padding rows [m, ld): identity, so the blocked inverse leaves them apart.
__global__ void k_dense_pad_identity(int m, int ld, double* __restrict__ D)
{
    const int i = m + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ld) D[(long)i * ld + i] = 1.0;
}

// ---------------------------------------------------- blocked Gauss-Jordan
// In-place inverse of the SPD ld x ld row-major matrix D, ld = 96 T, without
// pivoting (the convention of the domain inverses, MASPreconditioner.cu:
// 630-726), in 96 x 96 tiles.  Step k of T:
//   P = D_kk^-1 (gj_invert on one workgroup), C_i = D_ik (i != k)  k_gj_pivot
//   D_kj = P D_kj (j != k)                                          k_gj_row
//   D_ij -= C_i D_kj (i, j != k), D_ik = -C_i P, D_kk = P            k_gj_update
// which is the scalar Gauss-Jordan step applied to whole tiles.  A pivot that
// is not positive sets *nonpos.  Every product sums in a fixed order.
constexpr int kGjTile = 96;  // = 3B at B = 32: gj_invert<32>'s 16 x 16 grid of 6 x 6 tiles
constexpr int kGjChunk = 16; // inner-dimension chunk staged in LDS

// acc += X Y for 96 x 96 tiles X (row stride ldx) and Y (row stride ldy);
// thread (ty, tx) accumulates rows 6ty.., cols 6tx..
__device__ __forceinline__ void gj_tile_gemm(const double* __restrict__ X, int ldx, const double* __restrict__ Y,
                                             int ldy, double acc[6][6], double* __restrict__ sX,
                                             double* __restrict__ sY, int ty, int tx)
{
    for (int t0 = 0; t0 < kGjTile; t0 += kGjChunk) {
        for (int e = threadIdx.x; e < kGjTile * kGjChunk; e += blockDim.x) {
            const int row = e / kGjChunk, tt = e % kGjChunk;
            sX[tt * kGjTile + row] = X[(long)row * ldx + t0 + tt];
            const int tr = e / kGjTile, col = e % kGjTile;
            sY[tr * kGjTile + col] = Y[(long)(t0 + tr) * ldy + col];
        }
        __syncthreads();
#pragma unroll 4
        for (int tt = 0; tt < kGjChunk; ++tt) {
            double xr[6], yc[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                xr[q] = sX[tt * kGjTile + 6 * ty + q];
                yc[q] = sY[tt * kGjTile + 6 * tx + q];
            }
#pragma unroll
            for (int r = 0; r < 6; ++r)
#pragma unroll
                for (int c = 0; c < 6; ++c) acc[r][c] = fma(xr[r], yc[c], acc[r][c]);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_gj_pivot(const double* __restrict__ D, int ld, int k,
                                                  double* __restrict__ P, double* __restrict__ C,
                                                  int* __restrict__ nonpos)
{
    __shared__ double colbuf[2 * kGjTile], rowbuf[2 * kGjTile];
    const int i = blockIdx.x;
    const double* src = D + (long)i * kGjTile * ld + (long)k * kGjTile;
    if (i != k) {
        double* dst = C + (long)i * kGjTile * kGjTile;
        for (int e = threadIdx.x; e < kGjTile * kGjTile; e += blockDim.x)
            dst[e] = src[(long)(e / kGjTile) * ld + e % kGjTile];
        return;
    }
    const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
    double a[6][6];
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) a[r][c] = src[(long)(6 * ty + r) * ld + 6 * tx + c];
    gj_invert<32>(a, colbuf, rowbuf, ty, tx, nonpos);
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) P[(6 * ty + r) * kGjTile + 6 * tx + c] = a[r][c];
}

__global__ __launch_bounds__(256) void k_gj_row(double* __restrict__ D, int ld, int k, const double* __restrict__ P)
{
    __shared__ double sX[kGjChunk * kGjTile], sY[kGjChunk * kGjTile];
    const int j = blockIdx.x + (blockIdx.x >= k ? 1 : 0);
    const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
    double* t = D + (long)k * kGjTile * ld + (long)j * kGjTile;
    double acc[6][6] = {};
    gj_tile_gemm(P, kGjTile, t, ld, acc, sX, sY, ty, tx);
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) t[(long)(6 * ty + r) * ld + 6 * tx + c] = acc[r][c];
}

__global__ __launch_bounds__(256) void k_gj_update(double* __restrict__ D, int ld, int k, int T,
                                                   const double* __restrict__ P, const double* __restrict__ C)
{
    __shared__ double sX[kGjChunk * kGjTile], sY[kGjChunk * kGjTile];
    const int b = blockIdx.x;
    double* kk = D + (long)k * kGjTile * ld + (long)k * kGjTile;
    if (b == (T - 1) * T) {  // the pivot tile takes P
        for (int e = threadIdx.x; e < kGjTile * kGjTile; e += blockDim.x)
            kk[(long)(e / kGjTile) * ld + e % kGjTile] = P[e];
        return;
    }
    const int i = b / T + (b / T >= k ? 1 : 0), j = b % T;
    const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
    const double* Ci = C + (long)i * kGjTile * kGjTile;
    double* t = D + (long)i * kGjTile * ld + (long)j * kGjTile;
    double acc[6][6] = {};
    if (j == k) {
        gj_tile_gemm(Ci, kGjTile, P, kGjTile, acc, sX, sY, ty, tx);
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = 0; c < 6; ++c) t[(long)(6 * ty + r) * ld + 6 * tx + c] = -acc[r][c];
    } else {
        gj_tile_gemm(Ci, kGjTile, D + (long)k * kGjTile * ld + (long)j * kGjTile, ld, acc, sX, sY, ty, tx);
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                double* e = t + (long)(6 * ty + r) * ld + 6 * tx + c;
                *e = *e - acc[r][c];
            }
    }
}

int grid_for(long n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace

int mas_num_levels(int n, int B)
{
    int nLevel = 1;
    int levelSz = (n + B - 1) / B * B;
    while (levelSz > B) {
        levelSz /= B;
        nLevel++;
        levelSz = (levelSz + B - 1) / B * B;
    }
    nLevel = nLevel + 1;
    return nLevel > kMasMaxLevels ? kMasMaxLevels : nLevel;
}

Mas::Mas(Context* c, int bank) : ctx(c), B(bank)
{
    SGIPC_CHECK(bank == 16 || bank == 32, SGIPC_BAD_ARG, "bank_size must be 16 or 32");
}

void Mas::init_neighbor(int vn, int cof, int total_nbr, const uint32_t* start, const uint32_t* num,
                        const uint32_t* list)
{
    SGIPC_CHECK(vn >= 1, SGIPC_EMPTY, "MAS needs at least one vertex");
    SGIPC_CHECK(start && num && (list || total_nbr == 0), SGIPC_BAD_ARG, "null neighbour arrays");
    n = vn;
    coff = cof;
    L = mas_num_levels(n, B);
    hipStream_t s = ctx->stream;
    nbr_start_.reserve(n);
    nbr_num_.reserve(n);
    nbr_list_.reserve(total_nbr > 0 ? total_nbr : 1);
    SGIPC_HIP(hipMemcpyAsync(nbr_start_.get(), start, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    SGIPC_HIP(hipMemcpyAsync(nbr_num_.get(), num, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    if (total_nbr > 0)
        SGIPC_HIP(hipMemcpyAsync(nbr_list_.get(), list, sizeof(uint32_t) * total_nbr, hipMemcpyDeviceToDevice, s));
    map_.reserve((size_t)L * n);
    meta_.reserve(kMetaSize);
    // every level has at most n nodes: off_L <= L * (n + B)
    going_next_.reserve((size_t)L * (n + B));
    mask_.reserve(n);
    comp_rank_.reserve(n + 64);
    dom_count_.reserve(n / B + 2);
    dom_prefix_.reserve(n / B + 2);
    next_id_.reserve(n);
    row_ptr_.reserve(n + 1);
    built = false;
}

template <int BB>
void Mas::build_hierarchy_t(hipStream_t s, const int* pairs, int n_pairs)
{
    std::vector<int> hmeta(kMetaSize, 0);
    hmeta[kMetaCnt + 0] = n;
    hmeta[kMetaOff + 0] = 0;
    SGIPC_HIP(hipMemcpyAsync(meta_.get(), hmeta.data(), sizeof(int) * kMetaSize, hipMemcpyHostToDevice, s));
    const int gn = grid_for(n, 256);
    for (int l = 0; l < L; ++l) {
        const int* map_l = l == 0 ? nullptr : map_.get() + (size_t)l * n;
        if (l > 0) launch_fill<int>(s, reinterpret_cast<int*>(mask_.get()), 0, n);
        k_mask_fine<BB><<<gn, 256, 0, s>>>(l, n, map_l, nbr_start_.get(), nbr_num_.get(), nbr_list_.get(), mask_.get());
        SGIPC_LAUNCH_CHECK();
        if (n_pairs > 0) {
            const int gc = std::min(grid_for(n_pairs, 256), 128);
            k_mask_contact<BB><<<gc, 256, 0, s>>>(meta_.get(), l, n, coff, reinterpret_cast<const int4*>(pairs),
                                                  n_pairs, map_l, mask_.get());
            SGIPC_LAUNCH_CHECK();
        }
        // upper bound on the node count of level l is n, padded to whole blocks
        const int gpad = grid_for((long)((n + BB - 1) / BB) * BB, 256);
        k_closure<BB><<<gpad, 256, 0, s>>>(meta_.get(), l, mask_.get(), comp_rank_.get(), dom_count_.get());
        SGIPC_LAUNCH_CHECK();
        k_scan_domains<<<1, 1024, 0, s>>>(meta_.get(), l, BB, dom_count_.get(), dom_prefix_.get());
        SGIPC_LAUNCH_CHECK();
        k_assign<<<gn, 256, 0, s>>>(meta_.get(), l, BB, comp_rank_.get(), dom_prefix_.get(), next_id_.get(),
                                    going_next_.get());
        SGIPC_LAUNCH_CHECK();
        if (l + 1 < L) {
            k_map_next<<<gn, 256, 0, s>>>(n, map_l, next_id_.get(), map_.get() + (size_t)(l + 1) * n);
            SGIPC_LAUNCH_CHECK();
        }
    }
    SGIPC_HIP(hipMemcpyAsync(hmeta.data(), meta_.get(), sizeof(int) * kMetaSize, hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    for (int l = 0; l <= L; ++l) {
        h_cnt[l] = hmeta[kMetaCnt + l];
        h_off[l] = hmeta[kMetaOff + l];
    }
    total = h_off[L];
}

void Mas::build_hierarchy(hipStream_t s, const int* pairs, int n_pairs)
{
    if (B == 32) build_hierarchy_t<32>(s, pairs, n_pairs);
    else build_hierarchy_t<16>(s, pairs, n_pairs);
}

void Mas::prebuild_hierarchy(hipStream_t s, const int* pairs, int n_pairs)
{
    SGIPC_CHECK(n >= 1, SGIPC_EMPTY, "init_neighbor first");
    SGIPC_CHECK(n_pairs == 0 || pairs != nullptr, SGIPC_BAD_ARG, "null pairs");
    hier_ready_ = false;
    build_hierarchy(s, pairs, n_pairs);  // ends with a read-back: s is drained
    hier_ready_ = true;
    hier_pairs_ = pairs;
    hier_npairs_ = n_pairs;
}

template <int BB>
void Mas::assemble_invert_t(const double* vals, const int* rows, const int* cols, int nnz, int offset,
                            bool want_ap)
{
    hipStream_t s = ctx->stream;
    constexpr int nblk = BB * (BB + 1) / 2;
    const int ndom0 = h_off[1] / BB;
    const int ndom = total / BB;
    const int ndomc = ndom - ndom0;
    if (side_.s) side_.end(s);  // side-stream work of an earlier build that ended in an error
    mgal_.reserve((size_t)(ndomc > 0 ? ndomc : 1) * nblk * 9);
    if (ndomc > 0) launch_fill<double>(s, mgal_.get(), 0.0, (long)ndomc * nblk * 9);
    k_row_ptr<<<grid_for(n + 1, 256), 256, 0, s>>>(rows, nnz, offset, n, row_ptr_.get());
    SGIPC_LAUNCH_CHECK();
    // with a row offset the fine operator is the BCOO's suffix of rows >=
    // offset (rows sorted; an upper block of such a row has its column past
    // the offset too), renumbered from 0
    sub_e0_ = 0;
    sub_nnz_ = nnz;
    if (want_ap && offset > 0) {
        int e[2] = {0, 0};
        SGIPC_HIP(hipMemcpyAsync(&e[0], row_ptr_.get(), sizeof(int), hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipMemcpyAsync(&e[1], row_ptr_.get() + n, sizeof(int), hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipStreamSynchronize(s));
        SGIPC_CHECK(e[1] == nnz, SGIPC_BAD_INDEX, "blocks past the MAS's rows");
        sub_e0_ = e[0];
        sub_nnz_ = nnz - e[0];
        sub_rows_.reserve(sub_nnz_ > 0 ? sub_nnz_ : 1);
        sub_cols_.reserve(sub_nnz_ > 0 ? sub_nnz_ : 1);
        if (sub_nnz_ > 0) {
            k_shift_rc<<<grid_for(sub_nnz_, 256), 256, 0, s>>>(rows, cols, sub_e0_, sub_nnz_, offset,
                                                               sub_rows_.get(), sub_cols_.get());
            SGIPC_LAUNCH_CHECK();
        }
    }
    const double* fv = vals + 9 * sub_e0_;
    const int* fr = offset > 0 ? sub_rows_.get() : rows;
    const int* fc = offset > 0 ? sub_cols_.get() : cols;
    // A P of coarse mode 2 (needs the hierarchy and the fine matrix only) on
    // the side stream, beside the coarse levels; joined below
    if (want_ap && sub_nnz_ > 0) {
        side_.begin(s);
        build_ap(side_.s, fv, fr, fc, sub_nnz_);
    }
    // coarse operators A_l = P_l^T A_{l-1} P_l, level by level
    if (L > 1 && nnz > 0) {
        coarse_v_[0].reserve((size_t)nnz * 9);
        coarse_v_[1].reserve((size_t)nnz * 9);
        coarse_r_[0].reserve(nnz);
        coarse_r_[1].reserve(nnz);
        coarse_c_[0].reserve(nnz);
        coarse_c_[1].reserve(nnz);
        map_v_.reserve((size_t)nnz * 9);
        map_r_.reserve(nnz);
        map_c_.reserve(nnz);
        const double* pv = vals;
        const int* pr = rows;
        const int* pc = cols;
        int pn = nnz, poff = offset;
        for (int l = 1; l < L; ++l) {
            k_coarsen_map<<<grid_for(pn, 256), 256, 0, s>>>(pv, pr, pc, pn, poff, h_cnt[l - 1], going_next_.get(),
                                                            h_off[l - 1], h_off[l], map_v_.get(), map_r_.get(),
                                                            map_c_.get());
            SGIPC_LAUNCH_CHECK();
            const int q = l & 1;
            const int cn = conv_.convert(s, map_v_.get(), map_r_.get(), map_c_.get(), pn, h_cnt[l], coarse_v_[q].get(),
                                         coarse_r_[q].get(), coarse_c_[q].get(), nnz);
            k_extract_domains<BB><<<grid_for(cn, 256), 256, 0, s>>>(coarse_v_[q].get(), coarse_r_[q].get(),
                                                                     coarse_c_[q].get(), cn, h_off[l], ndom0,
                                                                     mgal_.get());
            SGIPC_LAUNCH_CHECK();
            if (l == 2 && coarse_mode >= 1) build_dense_level2(s, coarse_v_[q].get(), coarse_r_[q].get(),
                                                              coarse_c_[q].get(), cn);
            if (l == 1 && want_graph()) {
                graph_.build(s, coarse_v_[q].get(), coarse_r_[q].get(), coarse_c_[q].get(), cn, h_cnt[1]);
                g1d_.reserve(3L * h_cnt[1]);
                c2d_.reserve(3L * h_cnt[1]);
                g1f_.reserve(3L * h_cnt[1]);
                c2f_.reserve(3L * h_cnt[1]);
            }
            pv = coarse_v_[q].get();
            pr = coarse_r_[q].get();
            pc = coarse_c_[q].get();
            pn = cn;
            poff = 0;
        }
    }
    // the domain inverses (the level-2 Gauss-Jordan chain, a few workgroups
    // per launch, runs beside them on the side stream)
    constexpr int T = InvShape<BB>::T;
    if (precision == SGIPC_FP64) {
        minv64_.reserve((size_t)ndom * nblk * 9);
        k_inv_level0<BB, double><<<ndom0, T, 0, s>>>(n, vals, cols, row_ptr_.get(), offset, minv64_.get());
        SGIPC_LAUNCH_CHECK();
        if (ndomc > 0) {
            k_inv_coarse<BB, double><<<ndomc, T, 0, s>>>(mgal_.get(), ndom0, minv64_.get());
            SGIPC_LAUNCH_CHECK();
        }
    } else {
        minv32_.reserve((size_t)ndom * nblk * 9);
        k_inv_level0<BB, float><<<ndom0, T, 0, s>>>(n, vals, cols, row_ptr_.get(), offset, minv32_.get());
        SGIPC_LAUNCH_CHECK();
        if (ndomc > 0) {
            k_inv_coarse<BB, float><<<ndomc, T, 0, s>>>(mgal_.get(), ndom0, minv32_.get());
            SGIPC_LAUNCH_CHECK();
        }
        if (half_level0 && coarse_mode == 2 && ndom0 > 0) {
            minv16_.reserve((size_t)ndom0 * nblk * 9);
            msc_.reserve((size_t)ndom0 * BB * 3);
            k_half_level0<BB><<<ndom0, 256, 0, s>>>(minv32_.get(), minv16_.get(), msc_.get());
            SGIPC_LAUNCH_CHECK();
            half_ready_ = true;
        }
    }
    if (side_.s) side_.end(s);
    R64_.reserve((size_t)total * 3);
    Z64_.reserve((size_t)total * 3);
    R32_.reserve((size_t)total * 3);
    Z32_.reserve((size_t)total * 3);
}

// The dense level-2 inverse: the Galerkin operator A_2 scattered into a
// dense array padded to whole 96 x 96 tiles (identity on the padding), then
// the blocked Gauss-Jordan inverse on the build stream, 3 launches per tile
// row.  A pivot that is not positive means A_2 lost its definiteness to
// rounding: set_bcoo then returns SGIPC_NOT_SPD rather than running a cycle
// without its exact coarse solve.
void Mas::build_dense_level2(hipStream_t s, const double* v, const int* r, const int* c, int nb)
{
    dense_m_ = 0;
    dense_pending_ = false;
    const long m = 3L * h_cnt[2];
    if (m < 1 || m > kDenseMax || nb < 1) return;
    const int T = (int)((m + kGjTile - 1) / kGjTile);
    const int ld = T * kGjTile;
    dense_inv_.reserve((size_t)ld * ld);
    dense_piv_.reserve((size_t)kGjTile * kGjTile);
    dense_col_.reserve((size_t)T * kGjTile * kGjTile);
    dense_info_.reserve(1);
    launch_fill<int>(s, dense_info_.get(), 0, 1);
    launch_fill<double>(s, dense_inv_.get(), 0.0, (long)ld * ld);
    k_dense_scatter<<<grid_for(nb, 256), 256, 0, s>>>(v, r, c, nb, ld, dense_inv_.get());
    SGIPC_LAUNCH_CHECK();
    if (ld > m) {
        k_dense_pad_identity<<<grid_for(ld - m, 256), 256, 0, s>>>((int)m, ld, dense_inv_.get());
        SGIPC_LAUNCH_CHECK();
    }
    // the elimination runs on the side stream: its launches are small and
    // serial, so the level-0 domain inverses fill the chip meanwhile
    // (assemble_invert_t joins the streams after them)
    side_.begin(s);
    hipStream_t g = side_.s;
    for (int k = 0; k < T; ++k) {
        k_gj_pivot<<<T, 256, 0, g>>>(dense_inv_.get(), ld, k, dense_piv_.get(), dense_col_.get(),
                                     dense_info_.get());
        SGIPC_LAUNCH_CHECK();
        if (T > 1) {
            k_gj_row<<<T - 1, 256, 0, g>>>(dense_inv_.get(), ld, k, dense_piv_.get());
            SGIPC_LAUNCH_CHECK();
        }
        k_gj_update<<<(T - 1) * T + 1, 256, 0, g>>>(dense_inv_.get(), ld, k, T, dense_piv_.get(), dense_col_.get());
        SGIPC_LAUNCH_CHECK();
    }
    dense_pending_ = true;
    dense_m_pending_ = (int)m;
    dense_ld_ = ld;
}

void Mas::finish_dense_level2(hipStream_t s)
{
    if (!dense_pending_) return;
    dense_pending_ = false;
    int info = 0;
    SGIPC_HIP(hipMemcpyAsync(&info, dense_info_.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    SGIPC_CHECK(info == 0, SGIPC_NOT_SPD, "the level-2 Galerkin operator is not positive definite");
    dense_m_ = dense_m_pending_;
}

Mas::~Mas() = default;

void Mas::learned_sample(hipStream_t s, const double* r, double* ztmp, float* X, float* q)
{
    SGIPC_CHECK(built && graph_.n == h_cnt[1] && h_cnt[1] > 0, SGIPC_BAD_ARG, "MAS has no level-1 graph");
    Mpnn* keep = net_;
    net_ = nullptr;  // the plain preconditioner
    apply<double>(s, r, ztmp, nullptr, PcgStopDesc{});
    net_ = keep;
    const int c1 = h_cnt[1];
    const bool cyc = cycle_active();
    if (precision == SGIPC_FP64) {
        k_l1_features<double><<<grid_for(c1, 4), 256, 0, s>>>(
            nullptr, meta_.get(), L, going_next_.get(), c1, graph_.row_ptr.get(), graph_.col.get(), graph_.blk.get(),
            graph_.s.get(), R64_.get(), Z64_.get(), cyc ? g1d_.get() : nullptr, cyc ? c2d_.get() : nullptr, X);
    } else {
        k_l1_features<float><<<grid_for(c1, 4), 256, 0, s>>>(
            nullptr, meta_.get(), L, going_next_.get(), c1, graph_.row_ptr.get(), graph_.col.get(), graph_.blk.get(),
            graph_.s.get(), R32_.get(), Z32_.get(), cyc ? g1f_.get() : nullptr, cyc ? c2f_.get() : nullptr, X);
    }
    SGIPC_LAUNCH_CHECK();
    k_sample_normalise<<<1, 1024, 0, s>>>(3 * c1, X, q);
    SGIPC_LAUNCH_CHECK();
}

// Without a read-back: the entry count n_uniq stays on the device (the last
// element of the flags' inclusive scan) and the CSR arrays are sized for the
// upper bound 2 nnz, so the build can run on a side stream beside the
// coarse levels.
void Mas::build_ap(hipStream_t s, const double* vals, const int* rows, const int* cols, int nnz)
{
    SGIPC_CHECK(nnz < (1 << 30), SGIPC_CAPACITY, "A P entry count overflows int");
    const int m = 2 * nnz;
    int kb = 1;
    while (kb < 31 && (1L << kb) <= (long)n) ++kb;  // row n (the sentinel) fits too
    int cb = 1;
    while (cb < 31 && (1L << cb) < (long)h_cnt[1]) ++cb;  // level-1 columns 0 .. c1-1
    const int end_bit = kb + cb;
    apk_.reserve(m);
    apk_sorted_.reserve(m);
    api_.reserve(m);
    api_sorted_.reserve(m);
    apf_.reserve(m);
    apinc_.reserve(m);
    ap_ptr_.reserve(n + 1);
    ap_col_.reserve(m);
    ap_seg_.reserve((size_t)m + 1);
    ap_blk_.reserve(9L * m);
    size_t n1 = 0, n2 = 0;
    SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, n1, apk_.get(), apk_sorted_.get(), api_.get(), api_sorted_.get(),
                                        (size_t)m, 0, end_bit, s));
    SGIPC_HIP(rocprim::inclusive_scan(nullptr, n2, apf_.get(), apinc_.get(), (size_t)m, rocprim::plus<int>(), s));
    ap_temp_.reserve(n1 > n2 ? n1 : n2);
    k_ap_keys<<<grid_for(nnz, 256), 256, 0, s>>>(rows, cols, nnz, n, map_.get() + n, cb, apk_.get(), api_.get());
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(rocprim::radix_sort_pairs(ap_temp_.get(), n1, apk_.get(), apk_sorted_.get(), api_.get(),
                                        api_sorted_.get(), (size_t)m, 0, end_bit, s));
    k_ap_flags<<<grid_for(m, 256), 256, 0, s>>>(apk_sorted_.get(), m, n, cb, apf_.get());
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(rocprim::inclusive_scan(ap_temp_.get(), n2, apf_.get(), apinc_.get(), (size_t)m, rocprim::plus<int>(),
                                      s));
    k_ap_cols<<<grid_for(m, 256), 256, 0, s>>>(apk_sorted_.get(), apf_.get(), apinc_.get(), m, cb, ap_col_.get(),
                                               ap_seg_.get());
    SGIPC_LAUNCH_CHECK();
    k_ap_ptr<<<grid_for(n + 1, 256), 256, 0, s>>>(apk_sorted_.get(), apinc_.get(), m, n, cb, ap_ptr_.get(),
                                                  ap_seg_.get());
    SGIPC_LAUNCH_CHECK();
    k_ap_sum<<<grid_for(m, 256), 256, 0, s>>>(vals, api_sorted_.get(), ap_seg_.get(), apinc_.get(), m,
                                              ap_blk_.get());
    SGIPC_LAUNCH_CHECK();
    ap_nnz_ = m;  // an upper bound: A P is built
}

void Mas::init_matrix()
{
    SGIPC_CHECK(n >= 1, SGIPC_EMPTY, "init_neighbor first");
    hier_ready_ = false;
    build_hierarchy(ctx->stream, nullptr, 0);
}

void Mas::set_bcoo(const double* vals, const int* rows, const int* cols, int nnz, int offset,
                   const int* pairs, int n_pairs, int prec)
{
    SGIPC_CHECK(n >= 1, SGIPC_EMPTY, "init_neighbor first");
    SGIPC_CHECK(nnz >= 0, SGIPC_BAD_ARG, "negative nnz");
    SGIPC_CHECK(n_pairs == 0 || pairs != nullptr, SGIPC_BAD_ARG, "null pairs");
    auto t0 = std::chrono::steady_clock::now();
    precision = prec == SGIPC_FP64 ? SGIPC_FP64 : SGIPC_FP32_MAS;
    graph_.n = 0;
    dense_m_ = 0;
    half_ready_ = false;
    last_path_ = -1;
    dense_pending_ = false;  // a build that ended in an error left its check unread
    if (hier_ready_ && hier_pairs_ == pairs && hier_npairs_ == n_pairs) {
        hier_ready_ = false;  // built ahead by prebuild_hierarchy
    } else {
        hier_ready_ = false;
        build_hierarchy(ctx->stream, pairs, n_pairs);
    }
    ap_nnz_ = 0;
    // with a row offset the fine operator is always the MAS's own (the
    // caller's matrix also holds the rows before the offset)
    const bool want_ap = coarse_mode == 2 && L > 2 && nnz > 0 && (offset == 0 || build_own_fine);
    if (B == 32) assemble_invert_t<32>(vals, rows, cols, nnz, offset, want_ap);
    else assemble_invert_t<16>(vals, rows, cols, nnz, offset, want_ap);
    if (want_ap && build_own_fine && sub_nnz_ > 0) {
        const int* fr = offset > 0 ? sub_rows_.get() : rows;
        const int* fc = offset > 0 ? sub_cols_.get() : cols;
        own_fine_.build(ctx->stream, vals + 9 * sub_e0_, fr, fc, sub_nnz_, n, precision != SGIPC_FP64);
        fine_ = &own_fine_;
    }
    finish_dense_level2(ctx->stream);
    SGIPC_HIP(hipStreamSynchronize(ctx->stream));
    ms_build = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    built = true;
}

template <int BB, typename TV, typename TS>
int Mas::apply_t(hipStream_t s, const TV* r, TV* z, double* rz_part, const PcgStopDesc& sd)
{
    TS* R;
    TS* Z;
    const TS* minv;
    if constexpr (sizeof(TS) == 8) {
        R = R64_.get();
        Z = Z64_.get();
        minv = minv64_.get();
    } else {
        R = R32_.get();
        Z = Z32_.get();
        minv = minv32_.get();
    }
    const int* done = sd.ctl ? &sd.ctl->done : nullptr;
    const int npad = h_off[1];
    k_restrict0<BB, TV, TS><<<grid_for(npad, 256), 256, 0, s>>>(sd, meta_.get(), n, L, map_.get() + n, r, R);
    SGIPC_LAUNCH_CHECK();
    note_path(sizeof(TS) == 4);
    if (vcycle_active(sizeof(TS) == 4)) return finish_v_t<BB, TV, TS>(s, R, Z, minv, done, PcgStopDesc{}, r, z, rz_part);
    return finish_t<BB, TV, TS>(s, R, Z, minv, done, PcgStopDesc{}, r, z, rz_part);
}

template <int BB, typename TV, typename TS>
int Mas::finish_v_t(hipStream_t s, TS* R, TS* Z, const TS* minv, const int* done, const PcgStopDesc& sd2, const TV* r,
                    TV* z, double* rz_part)
{
    // R holds r at the level-0 slots (the identity slot map: slot u = row u)
    const int ndom0 = h_off[1] / BB;
    const int ndom01 = h_off[2] / BB;
    const int ndom = total / BB;
    const int npad = h_off[1];
    const int c1 = h_cnt[1];
    TS* G1;
    TS* C2;
    if constexpr (sizeof(TS) == 8) {
        G1 = g1d_.get();
        C2 = c2d_.get();
    } else {
        G1 = g1f_.get();
        C2 = c2f_.get();
    }
    const TS w = (TS)omega0;
    // the level-0 solves read the FP16 form of the inverse when it was built
    bool half = false;
    if constexpr (sizeof(TS) == 4) half = half_ready_;
    // pre-smoothing: Z0 = M0 r; then R0 = r1 = r - omega A Z0
    // FP16 level-0 solves.  The pre-smoothing takes one workgroup per domain
    // (measured 21.2 us on 8192 domains against 23.1 us persistent); the
    // post-smoothing, whose A P phase leaves the inverse loads time to land,
    // runs persistent with as many workgroups as are resident (29.1 against
    // 30.2 us).
    const uint4* m16 = reinterpret_cast<const uint4*>(minv16_.get());
    auto per_cu = [](const void* kern) {
        int k = 0;
        SGIPC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, kern, 256, 0));
        return std::max(k, 1);
    };
    auto grid_of = [&](int k) { return std::max(1, std::min(ndom0, k * ctx->num_cus)); };
    auto solve0 = [&]() {
        if constexpr (sizeof(TS) == 4) {
            if (half) {
                k_solve0_half<BB><<<std::max(1, ndom0), 256, 0, s>>>(done, m16, msc_.get(), R, Z, ndom0);
                return;
            }
        }
        k_local_solve<BB, TS><<<ndom0, 256, 0, s>>>(done, minv, R, Z, 0);
    };
    solve0();
    SGIPC_LAUNCH_CHECK();
    // and the level-1 sums of r1: in the SpMV's epilogue (halo form), else
    // by their own kernel
    const Level1Sums l1{map_.get() + n, npad, c1, h_off[2] - npad - c1};
    if (!fine_->template apply_level1_sums<TS>(s, -w, Z, TS(1), R, done, l1)) {
        fine_->template apply<TS>(s, -w, Z, TS(1), R, nullptr, done);
        k_restrict0<BB, TS, TS><<<grid_for(npad, 256), 256, 0, s>>>(PcgStopDesc{}, meta_.get(), n, L, map_.get() + n,
                                                                    R, R);
        SGIPC_LAUNCH_CHECK();
    }
    // the level-1 cycle on P^T r1: its correction is C2 + Z1 (Z1 = Z at the level-1 slots)
    const int g1w = grid_for(c1, 4);
    k_local_solve<BB, TS><<<ndom01 - ndom0, 256, 0, s>>>(done, minv, R, Z, ndom0);
    SGIPC_LAUNCH_CHECK();
    k_l1_residual<TS, false><<<g1w, 256, 0, s>>>(done, meta_.get(), c1, graph_.row_ptr.get(), graph_.col.get(),
                                                  graph_.blk.get(), graph_.s.get(), R, Z, G1, C2);
    SGIPC_LAUNCH_CHECK();
    auto restrict_up_level = [&](int l) {
        int c = h_cnt[l - 1];
        int pad = h_off[l + 1] - h_off[l] - h_cnt[l];
        int work = c > pad ? c : pad;
        k_restrict_up<TS><<<grid_for(work > 0 ? work : 1, 256), 256, 0, s>>>(
            done, meta_.get(), l, going_next_.get(), R, l == 2 ? sd2 : PcgStopDesc{});
        SGIPC_LAUNCH_CHECK();
    };
    int Lc = L;
    if (dense_m_ > 0) {
        restrict_up_level(2);
        k_dense_gemv<TS><<<dense_m_, 256, 0, s>>>(done, dense_m_, dense_ld_, dense_inv_.get(), R + 3L * h_off[2],
                                                                Z + 3L * h_off[2]);
        SGIPC_LAUNCH_CHECK();
        Lc = 3;
    } else {
        for (int l = 2; l < L; ++l) restrict_up_level(l);
        if (ndom > ndom01) {
            k_local_solve<BB, TS><<<ndom - ndom01, 256, 0, s>>>(done, minv, R, Z, ndom01);
            SGIPC_LAUNCH_CHECK();
        }
    }
    // the chained correction C2 formed inside the second residual
    k_l1_residual<TS, true, true><<<g1w, 256, 0, s>>>(done, meta_.get(), c1, graph_.row_ptr.get(), graph_.col.get(),
                                                       graph_.blk.get(), graph_.s.get(), R, Z, G1, C2,
                                                       going_next_.get(), Lc);
    SGIPC_LAUNCH_CHECK();
    using V4 = typename Vec4Of<TS>::T;
    c4_.reserve((size_t)c1);  // double4 elements: room for c1 float4 too
    V4* c4 = reinterpret_cast<V4*>(c4_.get());
    k_local_solve<BB, TS><<<ndom01 - ndom0, 256, 0, s>>>(done, minv, R, Z, ndom0, C2, c4, h_off[1], c1);
    SGIPC_LAUNCH_CHECK();
    // learned correction (flexible PCG), from the level-1 state of the cycle
    const float* corr = nullptr;
    if (learned_active()) {
        lX_.reserve(9L * c1);
        lchat_.reserve(3L * c1);
        lcorr_.reserve(3L * c1);
        k_l1_features<TS><<<grid_for(c1, 4), 256, 0, s>>>(done, meta_.get(), L, going_next_.get(), c1,
                                                           graph_.row_ptr.get(), graph_.col.get(), graph_.blk.get(),
                                                           graph_.s.get(), R, Z, G1, C2, lX_.get());
        SGIPC_LAUNCH_CHECK();
        net_->forward(s, graph_, lX_.get(), 1, lchat_.get(), lcorr_.get(), done);
        corr = lcorr_.get();
    }
    // post-smoothing of r2 = r1 - A P c, as one pass of solves over r + r2:
    // Z0 = M0 (r + r2), z = omega Z0 + P c
    auto solve0_post = [&]() {
        if constexpr (sizeof(TS) == 4) {
            if (half) {
                static const int k = per_cu(reinterpret_cast<const void*>(k_solve0_post_half<BB, TV>));
                k_solve0_post_half<BB, TV><<<grid_of(k), 256, 0, s>>>(
                    done, m16, msc_.get(), R, Z, n, ndom0, ap_ptr_.get(), ap_col_.get(), ap_blk_.get(), c4, r);
                return;
            }
        }
        if constexpr (sizeof(TS) == 8) {
            k_post_rhs<BB, TV, TS><<<ndom0, 256, 0, s>>>(done, R, n, ap_ptr_.get(), ap_col_.get(), ap_blk_.get(),
                                                          c4, r);
            SGIPC_LAUNCH_CHECK();
            k_local_solve<BB, TS><<<ndom0, 256, 0, s>>>(done, minv, R, Z, 0);
        } else {
            k_local_solve_post<BB, TV, TS><<<ndom0, 256, 0, s>>>(done, minv, R, Z, n, ap_ptr_.get(),
                                                                  ap_col_.get(), ap_blk_.get(), c4, r);
        }
    };
    solve0_post();
    SGIPC_LAUNCH_CHECK();
    const int g = vec_grid(n);
    k_prolong<TV, TS><<<g, 256, 0, s>>>(done, meta_.get(), n, L, map_.get(), Z, r, z, rz_part, C2, corr,
                                        static_cast<const TV*>(flex_Ap_), flex_zap_, w);
    SGIPC_LAUNCH_CHECK();
    return g;
}

template <int BB, typename TV, typename TS>
int Mas::finish_t(hipStream_t s, TS* R, TS* Z, const TS* minv, const int* done, const PcgStopDesc& sd2, const TV* r,
                  TV* z, double* rz_part)
{
    const int ndom0 = h_off[1] / BB;
    const int ndom01 = L > 1 ? h_off[2] / BB : ndom0;
    const int ndom = total / BB;
    const bool cyc = cycle_active();
    auto restrict_up_level = [&](int l) {
        int c = h_cnt[l - 1];
        int pad = h_off[l + 1] - h_off[l] - h_cnt[l];
        int work = c > pad ? c : pad;
        k_restrict_up<TS><<<grid_for(work > 0 ? work : 1, 256), 256, 0, s>>>(
            done, meta_.get(), l, going_next_.get(), R, l == 2 ? sd2 : PcgStopDesc{});
        SGIPC_LAUNCH_CHECK();
    };
    auto restrict_up = [&]() {
        for (int l = 2; l < L; ++l) restrict_up_level(l);
    };
    TS* G1 = nullptr;
    TS* C2 = nullptr;
    if (cyc) {
        if constexpr (sizeof(TS) == 8) {
            G1 = g1d_.get();
            C2 = c2d_.get();
        } else {
            G1 = g1f_.get();
            C2 = c2f_.get();
        }
        const int c1 = h_cnt[1];
        const int g1w = grid_for(c1, 4);  // one wave per level-1 node
        // pre-smoothing (level-1 domain solves, with the level-0 solves)
        k_local_solve<BB, TS><<<ndom01, 256, 0, s>>>(done, minv, R, Z, 0);
        SGIPC_LAUNCH_CHECK();
        k_l1_residual<TS, false><<<g1w, 256, 0, s>>>(done, meta_.get(), c1, graph_.row_ptr.get(), graph_.col.get(),
                                                      graph_.blk.get(), graph_.s.get(), R, Z, G1, C2);
        SGIPC_LAUNCH_CHECK();
        // coarse correction of the level-1 residual: the exact level-2 solve
        // (dense inverse) when it was built, else the additive levels >= 2
        int Lc = L;
        if (dense_m_ > 0) {
            restrict_up_level(2);
            k_dense_gemv<TS><<<dense_m_, 256, 0, s>>>(done, dense_m_, dense_ld_, dense_inv_.get(),
                                                                    R + 3L * h_off[2], Z + 3L * h_off[2]);
            SGIPC_LAUNCH_CHECK();
            Lc = 3;
        } else {
            restrict_up();
            if (ndom > ndom01) {
                k_local_solve<BB, TS><<<ndom - ndom01, 256, 0, s>>>(done, minv, R, Z, ndom01);
                SGIPC_LAUNCH_CHECK();
            }
        }
        k_l1_residual<TS, true, true><<<g1w, 256, 0, s>>>(done, meta_.get(), c1, graph_.row_ptr.get(),
                                                           graph_.col.get(), graph_.blk.get(), graph_.s.get(), R, Z,
                                                           G1, C2, going_next_.get(), Lc);
        SGIPC_LAUNCH_CHECK();
        // post-smoothing
        k_local_solve<BB, TS><<<ndom01 - ndom0, 256, 0, s>>>(done, minv, R, Z, ndom0);
        SGIPC_LAUNCH_CHECK();
    } else {
        restrict_up();
        k_local_solve<BB, TS><<<ndom, 256, 0, s>>>(done, minv, R, Z, 0);
        SGIPC_LAUNCH_CHECK();
    }
    const float* corr = nullptr;
    if (learned_active()) {
        const int c1 = h_cnt[1];
        lX_.reserve(9L * c1);
        lchat_.reserve(3L * c1);
        lcorr_.reserve(3L * c1);
        if (cyc) {
            k_l1_features<TS><<<grid_for(c1, 4), 256, 0, s>>>(done, meta_.get(), L, going_next_.get(), c1,
                                                               graph_.row_ptr.get(), graph_.col.get(), graph_.blk.get(),
                                                               graph_.s.get(), R, Z, G1, C2, lX_.get());
        } else {
            k_l1_features<TS><<<grid_for(c1, 4), 256, 0, s>>>(done, meta_.get(), L, going_next_.get(), c1,
                                                               graph_.row_ptr.get(), graph_.col.get(), graph_.blk.get(),
                                                               graph_.s.get(), R, Z, nullptr, nullptr, lX_.get());
        }
        SGIPC_LAUNCH_CHECK();
        net_->forward(s, graph_, lX_.get(), 1, lchat_.get(), lcorr_.get(), done);
        corr = lcorr_.get();
    }
    const int g = vec_grid(n);
    k_prolong<TV, TS><<<g, 256, 0, s>>>(done, meta_.get(), n, L, map_.get(), Z, r, z, rz_part, C2, corr,
                                        static_cast<const TV*>(flex_Ap_), flex_zap_, TS(1));
    SGIPC_LAUNCH_CHECK();
    return g;
}

template <int BB, typename TV, typename TS>
int Mas::apply_xr_t(hipStream_t s, const FusedXR<TV>& fx, TV* z, double* rz_part, int* n_rr)
{
    TS* R;
    TS* Z;
    const TS* minv;
    if constexpr (sizeof(TS) == 8) {
        R = R64_.get();
        Z = Z64_.get();
        minv = minv64_.get();
    } else {
        R = R32_.get();
        Z = Z32_.get();
        minv = minv32_.get();
    }
    const int* done = &fx.ctl->done;
    const int npad = h_off[1];
    const int g0 = grid_for(npad, 256);
    if (g0 > fx.rr_cap) return -1;
    if (vcycle_active(sizeof(TS) == 4))
        k_restrict0_xr<BB, TV, TS, false><<<g0, 256, 0, s>>>(fx, meta_.get(), n, L, map_.get() + n, R);
    else
        k_restrict0_xr<BB, TV, TS><<<g0, 256, 0, s>>>(fx, meta_.get(), n, L, map_.get() + n, R);
    SGIPC_LAUNCH_CHECK();
    *n_rr = g0;
    const PcgStopDesc sd{fx.ctl, fx.rz, fx.n_rz, fx.rr, g0};
    note_path(sizeof(TS) == 4);
    if (vcycle_active(sizeof(TS) == 4)) return finish_v_t<BB, TV, TS>(s, R, Z, minv, done, sd, fx.r, z, rz_part);
    return finish_t<BB, TV, TS>(s, R, Z, minv, done, sd, fx.r, z, rz_part);
}

template <typename TV>
int Mas::apply_update_xr(hipStream_t s, const FusedXR<TV>& fx, TV* z, double* rz_part, int* n_rr)
{
    SGIPC_CHECK(built, SGIPC_BAD_ARG, "MAS used before setPreconditioner");
    if (L <= 2 || fx.ctl == nullptr || fx.n3 != 3L * n) return -1;
    if (precision == SGIPC_FP64) {
        if (B == 32) return apply_xr_t<32, TV, double>(s, fx, z, rz_part, n_rr);
        return apply_xr_t<16, TV, double>(s, fx, z, rz_part, n_rr);
    }
    if (B == 32) return apply_xr_t<32, TV, float>(s, fx, z, rz_part, n_rr);
    return apply_xr_t<16, TV, float>(s, fx, z, rz_part, n_rr);
}
template int Mas::apply_update_xr<double>(hipStream_t, const FusedXR<double>&, double*, double*, int*);
template int Mas::apply_update_xr<float>(hipStream_t, const FusedXR<float>&, float*, double*, int*);

template <typename TV>
int Mas::apply(hipStream_t s, const TV* r, TV* z, double* rz_part, const PcgStopDesc& sd)
{
    SGIPC_CHECK(built, SGIPC_BAD_ARG, "MAS used before setPreconditioner");
    if (precision == SGIPC_FP64) {
        if (B == 32) return apply_t<32, TV, double>(s, r, z, rz_part, sd);
        return apply_t<16, TV, double>(s, r, z, rz_part, sd);
    }
    if (B == 32) return apply_t<32, TV, float>(s, r, z, rz_part, sd);
    return apply_t<16, TV, float>(s, r, z, rz_part, sd);
}
template int Mas::apply<double>(hipStream_t, const double*, double*, double*, const PcgStopDesc&);
template int Mas::apply<float>(hipStream_t, const float*, float*, double*, const PcgStopDesc&);

int Mas::levels(int* offsets, int* counts) const
{
    for (int l = 0; l <= L; ++l) offsets[l] = h_off[l];
    for (int l = 0; l < L; ++l) counts[l] = h_cnt[l];
    return L;
}

namespace {
__global__ void k_tables(int n, int L, const int* __restrict__ meta, const int* __restrict__ map,
                         int* __restrict__ coarse_table)
{
    int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    for (int l = 1; l < L; ++l)
        coarse_table[(long)u * (L - 1) + (l - 1)] = meta[kMetaOff + l] + map[(long)l * n + u];
}
__global__ void k_widen(const float* __restrict__ in, double* __restrict__ out, long cnt)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) out[i] = (double)in[i];
}
}  // namespace

void Mas::tables(int* coarse_table, int* going_next) const
{
    hipStream_t s = ctx->stream;
    if (coarse_table && L > 1) {
        k_tables<<<grid_for(n, 256), 256, 0, s>>>(n, L, meta_.get(), map_.get(), coarse_table);
        SGIPC_LAUNCH_CHECK();
    }
    if (going_next)
        SGIPC_HIP(hipMemcpyAsync(going_next, going_next_.get(), sizeof(int) * total, hipMemcpyDeviceToDevice, s));
    SGIPC_HIP(hipStreamSynchronize(s));
}

int Mas::level2_inverse(double* out) const
{
    if (dense_m_ <= 0 || out == nullptr) return dense_m_;
    hipStream_t s = ctx->stream;
    SGIPC_HIP(hipMemcpy2DAsync(out, sizeof(double) * dense_m_, dense_inv_.get(), sizeof(double) * dense_ld_,
                               sizeof(double) * dense_m_, dense_m_, hipMemcpyDeviceToDevice, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    return dense_m_;
}

void Mas::inverse_blocks(double* out) const
{
    hipStream_t s = ctx->stream;
    long cnt = (long)(total / B) * (B * (B + 1) / 2) * 9;
    if (precision == SGIPC_FP64) {
        SGIPC_HIP(hipMemcpyAsync(out, minv64_.get(), sizeof(double) * cnt, hipMemcpyDeviceToDevice, s));
    } else {
        k_widen<<<grid_for(cnt, 256), 256, 0, s>>>(minv32_.get(), out, cnt);
        SGIPC_LAUNCH_CHECK();
    }
    SGIPC_HIP(hipStreamSynchronize(s));
}

}  // namespace sgipc
