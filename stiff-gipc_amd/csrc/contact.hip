// Contact-pair generation and the barrier of every pair type, for gfx950.
//
// Reference: the LBVH self queries of StiffGIPC/mlbvh.cu (leaf boxes and
// Morton keys :1146-1225, radix tree :133-227,1247-1267, refit :1269-1296,
// point-face query :1306-1407, edge-edge query :1509-1632) with the
// distance-type narrow phase and pair encoding of _checkPTintersection /
// _checkEEintersection (:229-533, :535-1060, smooth = false), and the RANK 2 /
// NEWF branches of _calBarrierGradientAndHessian (StiffGIPC/GIPC.cu:3176-4984).
//
// Differences by design:
//   * the reference appends pairs with atomic counters, so their order (and
//     every MatIndex) changes from run to run; here each query runs twice --
//     a counting pass, a scan, a writing pass -- and the pairs are then sorted
//     by (query primitive, found primitive): point-face pairs by (surface
//     slot, face), then edge-edge pairs by (edge, edge).  MatIndex counts each
//     class in that order, so a detection is reproducible bit for bit;
//   * the refit reads the children's boxes with volatile loads after an
//     atomic arrival count, so a box written by another CU is never read from
//     a stale L1 line;
//   * a traversal deeper than its stack raises an error instead of writing
//     past it.
#include <algorithm>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "contact.hpp"
#include "pcg.hpp"
#include "wave_ops.hpp"

namespace sgipc {

namespace {

constexpr int kStack = 64;

inline int g256(long n) { return (int)((n + 255) / 256); }

// ------------------------------------------------------------ small vectors
struct V3 {
    double x, y, z;
};
__device__ __forceinline__ V3 ld3(const double* p, long i) { return V3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double det_cols(V3 c0, V3 c1, V3 c2) { return dot(c0, cross(c1, c2)); }

// squared distances (mlbvh.cu:229-271)
__device__ __forceinline__ double sq_pp(V3 a, V3 b)
{
    const V3 d = a - b;
    return dot(d, d);
}
__device__ __forceinline__ double sq_pe(V3 p, V3 a, V3 b)
{
    const V3 c = cross(a - p, b - p), e = b - a;
    return dot(c, c) / dot(e, e);
}
__device__ __forceinline__ double sq_pt(V3 p, V3 a, V3 b, V3 c)
{
    const V3 nv = cross(b - a, c - a);
    const double s = dot(p - a, nv);
    return s * s / dot(nv, nv);
}
__device__ __forceinline__ double sq_ee(V3 a0, V3 a1, V3 b0, V3 b1)
{
    const V3 nv = cross(a1 - a0, b1 - b0);
    const double s = dot(b0 - a0, nv);
    return s * s / dot(nv, nv);
}

// p in the frame (edge ab, in-plane normal of the edge, triangle normal)
__device__ __forceinline__ void edge_frame(V3 p, V3 a, V3 b, V3 nv, double& s, double& t)
{
    const V3 e = b - a, m = cross(e, nv), w = p - a;
    const double D = det_cols(e, m, nv);
    s = det_cols(w, m, nv) / D;
    t = det_cols(e, w, nv) / D;
}

// closest feature of triangle (t0, t1, t2) to p (mlbvh.cu:288-447)
__device__ int feature_pt(V3 p, V3 t0, V3 t1, V3 t2)
{
    const V3 nv = cross(t1 - t0, t2 - t0);
    double s0, u0, s1, u1, s2, u2;
    edge_frame(p, t0, t1, nv, s0, u0);
    if (s0 > 0.0 && s0 < 1.0 && u0 >= 0.0) return 3;
    edge_frame(p, t1, t2, nv, s1, u1);
    if (s1 > 0.0 && s1 < 1.0 && u1 >= 0.0) return 4;
    edge_frame(p, t2, t0, nv, s2, u2);
    if (s2 > 0.0 && s2 < 1.0 && u2 >= 0.0) return 5;
    if (s0 <= 0.0 && s2 >= 1.0) return 0;
    if (s1 <= 0.0 && s0 >= 1.0) return 1;
    if (s2 <= 0.0 && s1 >= 1.0) return 2;
    return 6;
}

// closest features of segments a0a1, b0b1 (mlbvh.cu:449-533)
__device__ int feature_ee(V3 a0, V3 a1, V3 b0, V3 b1)
{
    const V3 u = a1 - a0, v = b1 - b0, w = a0 - b0;
    const double a = dot(u, u), b = dot(u, v), c = dot(v, v), d = dot(u, w), e = dot(v, w);
    const double D = a * c - b * b;
    double tD = D, tN;
    int res = 8;
    const double sN = b * e - c * d;
    if (sN <= 0.0) {
        tN = e;
        tD = c;
        res = 2;
    } else if (sN >= D) {
        tN = e + b;
        tD = c;
        res = 5;
    } else {
        tN = a * e - b * d;
        const V3 uv = cross(u, v);
        if (tN > 0.0 && tN < tD && (dot(w, uv) == 0.0 || dot(uv, uv) < 1.0e-20 * a * c)) {
            tN = sN < D / 2 ? e : e + b;
            tD = c;
            res = sN < D / 2 ? 2 : 5;
        }
    }
    if (tN <= 0.0) return (-d <= 0.0) ? 0 : ((-d >= a) ? 3 : 6);
    if (tN >= tD) return ((-d + b) <= 0.0) ? 1 : (((-d + b) >= a) ? 4 : 7);
    return res;
}

struct Found {
    int4 pr;
    int cls;  // 2 PP, 3 PE, 4 PT/EE; 0: no contact
};

// _checkPTintersection (mlbvh.cu:535-651)
__device__ Found narrow_pt(const double* __restrict__ x, int p, int t0, int t1, int t2, double dsq)
{
    const V3 P = ld3(x, p), A = ld3(x, t0), B = ld3(x, t1), C = ld3(x, t2);
    Found f{make_int4(0, 0, 0, 0), 0};
    const int m = -p - 1;
    switch (feature_pt(P, A, B, C)) {
    case 0: if (sq_pp(P, A) < dsq) f = Found{make_int4(m, t0, -1, -1), 2}; break;
    case 1: if (sq_pp(P, B) < dsq) f = Found{make_int4(m, t1, -1, -1), 2}; break;
    case 2: if (sq_pp(P, C) < dsq) f = Found{make_int4(m, t2, -1, -1), 2}; break;
    case 3: if (sq_pe(P, A, B) < dsq) f = Found{make_int4(m, t0, t1, -1), 3}; break;
    case 4: if (sq_pe(P, B, C) < dsq) f = Found{make_int4(m, t1, t2, -1), 3}; break;
    case 5: if (sq_pe(P, C, A) < dsq) f = Found{make_int4(m, t2, t0, -1), 3}; break;
    default: if (sq_pt(P, A, B, C) < dsq) f = Found{make_int4(m, t0, t1, t2), 4}; break;
    }
    return f;
}

// _checkEEintersection, smooth = false (mlbvh.cu:685-1060): w of a
// point-point / point-edge pair is -obj - 2 for edges nearly parallel at rest
__device__ Found narrow_ee(const double* __restrict__ x, const double* __restrict__ rest, int a0, int a1, int b0, int b1,
                           int obj, double dsq)
{
    const V3 A0 = ld3(x, a0), A1 = ld3(x, a1), B0 = ld3(x, b0), B1 = ld3(x, b1);
    const int t = feature_ee(A0, A1, B0, B1);
    const V3 c = cross(A0 - A1, B0 - B1);
    const double ex = 1e-3 * sq_pp(ld3(rest, a0), ld3(rest, a1)) * sq_pp(ld3(rest, b0), ld3(rest, b1));
    const int we = dot(c, c) < ex ? -obj - 2 : -1;
    Found f{make_int4(0, 0, 0, 0), 0};
    switch (t) {
    case 0: if (sq_pp(A0, B0) < dsq) f = Found{make_int4(-a0 - 1, b0, -1, we), 2}; break;
    case 1: if (sq_pp(A0, B1) < dsq) f = Found{make_int4(-a0 - 1, b1, -1, we), 2}; break;
    case 2: if (sq_pe(A0, B0, B1) < dsq) f = Found{make_int4(-a0 - 1, b0, b1, we), 3}; break;
    case 3: if (sq_pp(A1, B0) < dsq) f = Found{make_int4(-a1 - 1, b0, -1, we), 2}; break;
    case 4: if (sq_pp(A1, B1) < dsq) f = Found{make_int4(-a1 - 1, b1, -1, we), 2}; break;
    case 5: if (sq_pe(A1, B0, B1) < dsq) f = Found{make_int4(-a1 - 1, b0, b1, we), 3}; break;
    case 6: if (sq_pe(B0, A0, A1) < dsq) f = Found{make_int4(-b0 - 1, a0, a1, we), 3}; break;
    case 7: if (sq_pe(B1, A0, A1) < dsq) f = Found{make_int4(-b1 - 1, a0, a1, we), 3}; break;
    default: if (sq_ee(A0, A1, B0, B1) < dsq) f = Found{make_int4(a0, a1, b0, b1), 4}; break;
    }
    return f;
}

// ------------------------------------------------------------ BVH build
// with dir: the box also covers x - alpha dir (swept boxes of the CCD queries,
// mlbvh.cu:1171-1204)
__global__ void k_prim_boxes(int n, const double* __restrict__ x, const uint32_t* __restrict__ prims, int k,
                             const double* __restrict__ dir, double alpha, double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double lo[3] = {1e32, 1e32, 1e32}, hi[3] = {-1e32, -1e32, -1e32};
    for (int a = 0; a < k; ++a) {
        const long v = prims[(long)k * i + a];
        for (int d = 0; d < 3; ++d) {
            const double c = x[3 * v + d];
            lo[d] = fmin(lo[d], c);
            hi[d] = fmax(hi[d], c);
            if (dir != nullptr) {
                const double e = c - dir[3 * v + d] * alpha;
                lo[d] = fmin(lo[d], e);
                hi[d] = fmax(hi[d], e);
            }
        }
    }
    for (int d = 0; d < 3; ++d) {
        out[6L * i + d] = lo[d];
        out[6L * i + 3 + d] = hi[d];
    }
}

// union of n boxes -> out[blockIdx.x]
__global__ __launch_bounds__(256) void k_box_union(int n, const double* __restrict__ in, double* __restrict__ out)
{
    __shared__ double sm[6][4];
    double b[6] = {1e32, 1e32, 1e32, -1e32, -1e32, -1e32};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int d = 0; d < 3; ++d) {
            b[d] = fmin(b[d], in[6L * i + d]);
            b[3 + d] = fmax(b[3 + d], in[6L * i + 3 + d]);
        }
    for (int d = 0; d < 6; ++d) {
        double v = b[d];
        for (int off = 32; off >= 1; off >>= 1) {
            const double o = __shfl_xor(v, off, 64);
            v = d < 3 ? fmin(v, o) : fmax(v, o);
        }
        b[d] = v;
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int d = 0; d < 6; ++d) sm[d][w] = b[d];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int d = threadIdx.x;
        double v = sm[d][0];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) v = d < 3 ? fmin(v, sm[d][q]) : fmax(v, sm[d][q]);
        out[6L * blockIdx.x + d] = v;
    }
}

// 10 bits per axis, interleaved (x highest)
__device__ __forceinline__ uint32_t spread10(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// key = Morton code of the box centre in the scene box (<< 32) | primitive
__global__ void k_morton(int n, const double* __restrict__ lb, const double* __restrict__ scene,
                         uint64_t* __restrict__ keys, int* __restrict__ idx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t q[3];
    for (int d = 0; d < 3; ++d) {
        const double ext = scene[3 + d] - scene[d];
        const double c = 0.5 * (lb[6L * i + d] + lb[6L * i + 3 + d]) - scene[d];
        double u = ext > 0.0 ? c / ext * 1024.0 : 0.0;
        u = fmin(fmax(u, 0.0), 1023.0);
        q[d] = (uint32_t)u;
    }
    const uint32_t mc = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    keys[i] = ((uint64_t)mc << 32) | (uint32_t)i;
    idx[i] = i;
}

__device__ __forceinline__ int common_prefix(const uint64_t* __restrict__ k, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    return __clzll(k[i] ^ k[j]);
}

// radix tree over the sorted keys (Karras 2012): internal node i covers a
// contiguous run of leaves that starts or ends at i; its split is where the
// keys' common prefix first grows
__global__ void k_radix_tree(int n, const uint64_t* __restrict__ k, int* __restrict__ child, int* __restrict__ parent)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int dir = common_prefix(k, n, i, i + 1) > common_prefix(k, n, i, i - 1) ? 1 : -1;
    const int pmin = common_prefix(k, n, i, i - dir);
    int lmax = 2;
    while (common_prefix(k, n, i, i + lmax * dir) > pmin) lmax <<= 1;
    int len = 0;
    for (int t = lmax >> 1; t > 0; t >>= 1)
        if (common_prefix(k, n, i, i + (len + t) * dir) > pmin) len += t;
    const int j = i + len * dir;
    const int pnode = common_prefix(k, n, i, j);
    int s = 0, t = len;
    do {
        t = (t + 1) >> 1;
        if (common_prefix(k, n, i, i + (s + t) * dir) > pnode) s += t;
    } while (t > 1);
    const int split = i + s * dir + (dir < 0 ? -1 : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const int left = (lo == split) ? n - 1 + split : split;
    const int right = (hi == split + 1) ? n - 1 + split + 1 : split + 1;
    child[2 * i] = left;
    child[2 * i + 1] = right;
    parent[left] = i;
    parent[right] = i;
}

// leaves take their primitive's box; the second child to arrive at a node
// forms its box and climbs on
__global__ void k_refit(int n, const int* __restrict__ sidx, const double* __restrict__ lb, double* box,
                        const int* __restrict__ child, const int* __restrict__ parent, int* flag)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int node = n - 1 + s;
    const long src = sidx[s];
    for (int d = 0; d < 6; ++d) box[6L * node + d] = lb[6 * src + d];
    __threadfence();
    int p = parent[node];
    while (p >= 0) {
        if (atomicAdd(flag + p, 1) == 0) return;
        __threadfence();
        const int c0 = child[2 * p], c1 = child[2 * p + 1];
        const volatile double* vb = box;
        double b[6];
        for (int d = 0; d < 3; ++d) {
            b[d] = fmin(vb[6L * c0 + d], vb[6L * c1 + d]);
            b[3 + d] = fmax(vb[6L * c0 + 3 + d], vb[6L * c1 + 3 + d]);
        }
        for (int d = 0; d < 6; ++d) box[6L * p + d] = b[d];
        __threadfence();
        p = parent[p];
    }
}

__device__ __forceinline__ bool boxes_overlap(const double* q, const double* __restrict__ b, double gap)
{
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if ((b[d] - q[3 + d]) >= gap || (q[d] - b[3 + d]) >= gap) return false;
    return true;
}

// Depth-first traversal from the root; visit(prim) for every leaf whose box
// overlaps q within gap.  Returns false when the stack would overflow.
template <typename F>
__device__ bool traverse(int n, const double* __restrict__ box, const int* __restrict__ child,
                         const int* __restrict__ sidx, const double* q, double gap, F&& visit)
{
    if (n == 1) {
        if (boxes_overlap(q, box, gap)) visit(sidx[0]);
        return true;
    }
    int stack[kStack];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const int node = stack[--sp];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int ch = child[2 * node + c];
            if (!boxes_overlap(q, box + 6L * ch, gap)) continue;
            if (ch >= n - 1) {
                visit(sidx[ch - (n - 1)]);
            } else {
                if (sp == kStack) return false;
                stack[sp++] = ch;
            }
        }
    }
    return true;
}

__device__ __forceinline__ bool all_fixed(const int* __restrict__ bt, int a, int b, int c, int d)
{
    return bt != nullptr && bt[a] >= 2 && bt[b] >= 2 && bt[c] >= 2 && bt[d] >= 2;
}

// point-face query (mlbvh.cu:1306-1407), one thread per surface vertex
template <bool kWrite>
__global__ __launch_bounds__(256) void k_query_vf(int n_surf, const uint32_t* __restrict__ surf,
                                                  const double* __restrict__ x, const uint32_t* __restrict__ faces,
                                                  int nf, const double* __restrict__ box,
                                                  const int* __restrict__ child, const int* __restrict__ sidx,
                                                  const int* __restrict__ body, const int* __restrict__ bt,
                                                  double dsq, int* __restrict__ cnt, const int* __restrict__ off,
                                                  uint64_t* __restrict__ rkey, int4* __restrict__ rpair,
                                                  int* __restrict__ rcls, int* __restrict__ err)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_surf) return;
    const int p = (int)surf[q];
    const V3 P = ld3(x, p);
    const double qb[6] = {P.x, P.y, P.z, P.x, P.y, P.z};
    const double gap = sqrt(dsq);
    int m = 0;
    const long base = kWrite ? off[q] : 0;
    const bool ok = traverse(nf, box, child, sidx, qb, gap, [&](int f) {
        const int t0 = (int)faces[3L * f], t1 = (int)faces[3L * f + 1], t2 = (int)faces[3L * f + 2];
        if (body != nullptr && body[p] == body[t0] && body[p] != -1) return;
        if (p == t0 || p == t1 || p == t2) return;
        if (all_fixed(bt, p, t0, t1, t2)) return;
        const Found fd = narrow_pt(x, p, t0, t1, t2, dsq);
        if (fd.cls == 0) return;
        if (kWrite) {
            rkey[base + m] = ((uint64_t)(uint32_t)q << 32) | (uint32_t)f;
            rpair[base + m] = fd.pr;
            rcls[base + m] = fd.cls;
        }
        ++m;
    });
    if (!ok) atomicOr(err, 1);
    if (!kWrite) cnt[q] = m;
}

// edge-edge query (mlbvh.cu:1509-1632), one thread per edge, each pair once
// (found edge > query edge)
template <bool kWrite>
__global__ __launch_bounds__(256) void k_query_ee(int ne, const uint32_t* __restrict__ edges,
                                                  const double* __restrict__ x, const double* __restrict__ rest,
                                                  const double* __restrict__ box, const int* __restrict__ child,
                                                  const int* __restrict__ sidx, const double* __restrict__ lb,
                                                  const int* __restrict__ body, const int* __restrict__ bt,
                                                  double dsq, int* __restrict__ cnt, const int* __restrict__ off,
                                                  uint64_t* __restrict__ rkey, int4* __restrict__ rpair,
                                                  int* __restrict__ rcls, int* __restrict__ err)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int a0 = (int)edges[2L * e], a1 = (int)edges[2L * e + 1];
    double qb[6];
    for (int d = 0; d < 6; ++d) qb[d] = lb[6L * e + d];
    const double gap = sqrt(dsq);
    int m = 0;
    const long base = kWrite ? off[e] : 0;
    const bool ok = traverse(ne, box, child, sidx, qb, gap, [&](int o) {
        if (o <= e) return;
        const int b0 = (int)edges[2L * o], b1 = (int)edges[2L * o + 1];
        if (body != nullptr && body[a0] == body[b0] && body[a0] != -1) return;
        if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return;
        if (all_fixed(bt, a0, a1, b0, b1)) return;
        const Found fd = narrow_ee(x, rest, a0, a1, b0, b1, o, dsq);
        if (fd.cls == 0) return;
        if (kWrite) {
            rkey[base + m] = ((uint64_t)(uint32_t)e << 32) | (uint32_t)o;
            rpair[base + m] = fd.pr;
            rcls[base + m] = fd.cls;
        }
        ++m;
    });
    if (!ok) atomicOr(err, 1);
    if (!kWrite) cnt[e] = m;
}

__global__ void k_iota(int n, int* __restrict__ a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// pairs in sorted order and one flag per class (2, 3, 4) for the MatIndex scans
__global__ void k_emit(int n, const int* __restrict__ order, const int4* __restrict__ rpair, const int* __restrict__ rcls,
                       int4* __restrict__ pairs, int* __restrict__ flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = order[i];
    pairs[i] = rpair[r];
    const int c = rcls[r];
    flags[i] = c == 2;
    flags[(long)n + i] = c == 3;
    flags[2L * n + i] = c == 4;
}

__global__ void k_mat_index(int n, const int* __restrict__ rcls, const int* __restrict__ order,
                            const int* __restrict__ scan, int* __restrict__ mat_index)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = rcls[order[i]] - 2;
    mat_index[i] = scan[(long)c * n + i];
}

// ------------------------------------------------------------ barrier
__device__ __forceinline__ double barrier_lam0(double I5, double dh, double kappa)
{
    const double L = log(I5);
    return -(4.0 * kappa * dh * dh *
             (4.0 * I5 + L - 3.0 * I5 * I5 * L * L + 6.0 * I5 * L - 2.0 * I5 * I5 + I5 * L * L - 7.0 * I5 * I5 * L - 2.0)) /
           I5;
}

// One thread per pair: the gradient of dist / sqrt(dHat) in closest-point
// form (point-point, point-line foot, point-plane foot with barycentric
// weights, line-line closest points), the reference's squared distance,
// then the rank-1 Hessian lambda0 g g^T and the gradient g s pk.
__global__ __launch_bounds__(128) void k_barrier_pairs(int np, const double* __restrict__ x,
                                                       const int4* __restrict__ pairs,
                                                       const int* __restrict__ mat_index, int n4, int n3, double dh,
                                                       double kappa, int hess_offset, double* __restrict__ grad,
                                                       double* __restrict__ tvals, int* __restrict__ trows,
                                                       int* __restrict__ tcols, int* __restrict__ bad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int4 pr = pairs[p];
    int vid[4] = {0, 0, 0, 0};
    int k;
    if (pr.x >= 0) {
        if (pr.w < 0) {  // smooth-encoded edge-edge pair: not produced by the detection
            atomicOr(bad, 1);
            return;
        }
        vid[0] = pr.x; vid[1] = pr.y; vid[2] = pr.z; vid[3] = pr.w;
        k = 5;
    } else {
        vid[0] = -pr.x - 1;
        vid[1] = pr.y;
        if (pr.y < 0) {
            atomicOr(bad, 1);
            return;
        }
        if (pr.z < 0) {
            k = 2;
        } else if (pr.w < 0) {
            vid[2] = pr.z;
            k = 3;
        } else {
            vid[2] = pr.z;
            vid[3] = pr.w;
            k = 4;
        }
    }
    const int nv = k == 5 ? 4 : k;
    V3 P[4];
    for (int a = 0; a < nv; ++a) P[a] = ld3(x, vid[a]);
    V3 r;
    double coef[4] = {0.0, 0.0, 0.0, 0.0};
    double d2;
    if (k == 2) {
        r = P[0] - P[1];
        coef[0] = 1.0;
        coef[1] = -1.0;
        d2 = sq_pp(P[0], P[1]);
    } else if (k == 3) {
        const V3 e = P[2] - P[1], u = P[0] - P[1];
        const double t = dot(u, e) / dot(e, e);
        r = V3{u.x - t * e.x, u.y - t * e.y, u.z - t * e.z};
        coef[0] = 1.0;
        coef[1] = -(1.0 - t);
        coef[2] = -t;
        d2 = sq_pe(P[0], P[1], P[2]);
    } else if (k == 4) {
        const V3 e1 = P[2] - P[1], e2 = P[3] - P[1], w = P[0] - P[1], nrm = cross(e1, e2);
        const double nn = dot(nrm, nrm);
        const double b1 = dot(cross(w, e2), nrm) / nn, b2 = dot(cross(e1, w), nrm) / nn;
        const double s = dot(w, nrm) / nn;
        r = V3{s * nrm.x, s * nrm.y, s * nrm.z};
        coef[0] = 1.0;
        coef[1] = -(1.0 - b1 - b2);
        coef[2] = -b1;
        coef[3] = -b2;
        d2 = sq_pt(P[0], P[1], P[2], P[3]);
    } else {
        const V3 ea = P[1] - P[0], eb = P[3] - P[2], w0 = P[0] - P[2];
        const double a = dot(ea, ea), b = dot(ea, eb), c = dot(eb, eb), dd = dot(ea, w0), e = dot(eb, w0);
        const double D = a * c - b * b;
        const double s = (b * e - c * dd) / D, t = (a * e - b * dd) / D;
        r = V3{w0.x + s * ea.x - t * eb.x, w0.y + s * ea.y - t * eb.y, w0.z + s * ea.z - t * eb.z};
        coef[0] = 1.0 - s;
        coef[1] = s;
        coef[2] = -(1.0 - t);
        coef[3] = -t;
        d2 = sq_ee(P[0], P[1], P[2], P[3]);
    }
    const double sq = sqrt(dh);
    const double rl = sqrt(dot(r, r));
    const double ux = r.x / rl / sq, uy = r.y / rl / sq, uz = r.z / rl / sq;
    double g[12];
    for (int a = 0; a < 4; ++a) {
        g[3 * a] = coef[a] * ux;
        g[3 * a + 1] = coef[a] * uy;
        g[3 * a + 2] = coef[a] * uz;
    }
    const double I5 = d2 / dh;
    const double lam0 = barrier_lam0(d2 < 1e-6 * dh ? 1e-6 : I5, dh, kappa);
    if (grad != nullptr) {
        const double L = log(I5);
        const double pk = 2.0 * (2.0 * kappa * dh * dh * L * (I5 - 1.0) * (I5 + I5 * L - 1.0)) / I5;
        const double sc = sqrt(d2) / sq * pk;
        for (int a = 0; a < nv; ++a)
            for (int d = 0; d < 3; ++d) atomic_add(grad + 3L * vid[a] + d, g[3 * a + d] * sc);
    }
    const int mi = mat_index[p];
    const long base = nv == 4 ? 16L * mi : (nv == 3 ? 16L * n4 + 9L * mi : 16L * n4 + 9L * n3 + 4L * mi);
    for (int ii = 0; ii < nv; ++ii)
        for (int jj = 0; jj < nv; ++jj) {
            const long kk = base + ii * nv + jj;
            const int row = vid[ii] + hess_offset, col = vid[jj] + hess_offset;
            double* dst = tvals + 9 * kk;
            if (row <= col) {
                trows[kk] = row;
                tcols[kk] = col;
                for (int rr = 0; rr < 3; ++rr)
                    for (int cc = 0; cc < 3; ++cc) dst[cc * 3 + rr] = lam0 * g[3 * ii + rr] * g[3 * jj + cc];
            } else {
                trows[kk] = col;
                tcols[kk] = row;
                for (int e = 0; e < 9; ++e) dst[e] = 0.0;
            }
        }
}


// ------------------------------------------------------------ feasible step
// distance of the closest features (ACCD.cu:317-371)
__device__ double pt_dist_any(V3 p, V3 a, V3 b, V3 c)
{
    switch (feature_pt(p, a, b, c)) {
    case 0: return sq_pp(p, a);
    case 1: return sq_pp(p, b);
    case 2: return sq_pp(p, c);
    case 3: return sq_pe(p, a, b);
    case 4: return sq_pe(p, b, c);
    case 5: return sq_pe(p, c, a);
    default: return sq_pt(p, a, b, c);
    }
}
__device__ double ee_dist_any(V3 a0, V3 a1, V3 b0, V3 b1)
{
    switch (feature_ee(a0, a1, b0, b1)) {
    case 0: return sq_pp(a0, b0);
    case 1: return sq_pp(a0, b1);
    case 2: return sq_pe(a0, b0, b1);
    case 3: return sq_pp(a1, b0);
    case 4: return sq_pp(a1, b1);
    case 5: return sq_pe(a1, b0, b1);
    case 6: return sq_pe(b0, a0, a1);
    case 7: return sq_pe(b1, a0, a1);
    default: return sq_ee(a0, a1, b0, b1);
    }
}

// Additive CCD (ACCD.cu:373-534): conservative advancement of the four points
// moving by D (centred), with eta = 1 - slackness and zero thickness; the time
// of contact in [0, 1], at most 50000 steps.
__device__ double accd_toc(V3 P[4], V3 D[4], bool ee, double eta)
{
    const V3 mov{-0.25 * (D[0].x + D[1].x + D[2].x + D[3].x), -0.25 * (D[0].y + D[1].y + D[2].y + D[3].y),
                 -0.25 * (D[0].z + D[1].z + D[2].z + D[3].z)};
    for (int a = 0; a < 4; ++a) D[a] = V3{D[a].x + mov.x, D[a].y + mov.y, D[a].z + mov.z};
    const double maxd = ee ? sqrt(fmax(dot(D[0], D[0]), dot(D[1], D[1]))) + sqrt(fmax(dot(D[2], D[2]), dot(D[3], D[3])))
                           : sqrt(dot(D[0], D[0])) +
                                 sqrt(fmax(dot(D[1], D[1]), fmax(dot(D[2], D[2]), dot(D[3], D[3]))));
    if (maxd == 0.0) return 1.0;
    auto dist2 = [&]() {
        if (!ee) return pt_dist_any(P[0], P[1], P[2], P[3]);
        double d2 = ee_dist_any(P[0], P[1], P[2], P[3]);
        if (d2 <= 0.0)
            d2 = fmin(fmin(sq_pp(P[0], P[2]), sq_pp(P[0], P[3])), fmin(sq_pp(P[1], P[2]), sq_pp(P[1], P[3])));
        return d2;
    };
    double d2 = dist2(), dc = sqrt(d2);
    const double gap = eta * d2 / dc;
    double toc = 0.0;
    for (int count = 1;; ++count) {
        if (count > 50000) return toc;
        const double tlb = (1.0 - eta) * d2 / (dc * maxd);
        for (int a = 0; a < 4; ++a) P[a] = V3{P[a].x + D[a].x * tlb, P[a].y + D[a].y * tlb, P[a].z + D[a].z * tlb};
        d2 = dist2();
        dc = sqrt(d2);
        if (toc != 0.0 && d2 / dc < gap) break;
        toc += tlb;
        if (toc > 1.0) return 1.0;
    }
    return toc;
}

// point-face candidates of the swept boxes (mlbvh.cu:1408-1508 with
// _checkPTintersection_fullCCD, :653-683): per surface vertex the largest
// 1 / toc over its candidates
__global__ __launch_bounds__(256) void k_ccd_vf(int n_surf, const uint32_t* __restrict__ surf,
                                                const double* __restrict__ x, const double* __restrict__ dir,
                                                double alpha, const uint32_t* __restrict__ faces, int nf,
                                                const double* __restrict__ box, const int* __restrict__ child,
                                                const int* __restrict__ sidx, const int* __restrict__ body,
                                                const int* __restrict__ bt, double dsq, double eta,
                                                double* __restrict__ inv_toc, int* __restrict__ cnt,
                                                int* __restrict__ err)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_surf) return;
    const int p = (int)surf[q];
    const V3 P = ld3(x, p), Pd = ld3(dir, p);
    const double qb[6] = {fmin(P.x, P.x - Pd.x * alpha), fmin(P.y, P.y - Pd.y * alpha), fmin(P.z, P.z - Pd.z * alpha),
                          fmax(P.x, P.x - Pd.x * alpha), fmax(P.y, P.y - Pd.y * alpha), fmax(P.z, P.z - Pd.z * alpha)};
    double best = 1.0;
    int m = 0;
    const bool ok = traverse(nf, box, child, sidx, qb, sqrt(dsq), [&](int f) {
        const int t0 = (int)faces[3L * f], t1 = (int)faces[3L * f + 1], t2 = (int)faces[3L * f + 2];
        if (body != nullptr && body[p] == body[t0] && body[p] != -1) return;
        if (p == t0 || p == t1 || p == t2) return;
        if (all_fixed(bt, p, t0, t1, t2)) return;
        const V3 A = ld3(x, t0), B = ld3(x, t1), C = ld3(x, t2);
        if (feature_pt(P, A, B, C) == 6 && dot(cross(B - A, C - A), P - A) < 0.0) return;
        ++m;
        V3 Q[4] = {P, A, B, C};
        const V3 da = ld3(dir, t0), db = ld3(dir, t1), dc = ld3(dir, t2);
        V3 D[4] = {V3{-Pd.x, -Pd.y, -Pd.z}, V3{-da.x, -da.y, -da.z}, V3{-db.x, -db.y, -db.z}, V3{-dc.x, -dc.y, -dc.z}};
        best = fmax(best, 1.0 / accd_toc(Q, D, false, eta));
    });
    if (!ok) atomicOr(err, 1);
    inv_toc[q] = best;
    cnt[q] = m;
}

// edge-edge candidates of the swept boxes (mlbvh.cu:1634-1774)
__global__ __launch_bounds__(256) void k_ccd_ee(int ne, const uint32_t* __restrict__ edges,
                                                const double* __restrict__ x, const double* __restrict__ dir,
                                                const double* __restrict__ box, const int* __restrict__ child,
                                                const int* __restrict__ sidx, const double* __restrict__ lb,
                                                const int* __restrict__ body, const int* __restrict__ bt,
                                                double dsq, double eta, double* __restrict__ inv_toc,
                                                int* __restrict__ cnt, int* __restrict__ err)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int a0 = (int)edges[2L * e], a1 = (int)edges[2L * e + 1];
    double qb[6];
    for (int d = 0; d < 6; ++d) qb[d] = lb[6L * e + d];
    double best = 1.0;
    int m = 0;
    const bool ok = traverse(ne, box, child, sidx, qb, sqrt(dsq), [&](int o) {
        if (o <= e) return;
        const int b0 = (int)edges[2L * o], b1 = (int)edges[2L * o + 1];
        if (body != nullptr && body[a0] == body[b0] && body[a0] != -1) return;
        if (all_fixed(bt, a0, a1, b0, b1)) return;
        if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return;
        ++m;
        V3 Q[4] = {ld3(x, a0), ld3(x, a1), ld3(x, b0), ld3(x, b1)};
        const int id[4] = {a0, a1, b0, b1};
        V3 D[4];
        for (int a = 0; a < 4; ++a) {
            const V3 dv = ld3(dir, id[a]);
            D[a] = V3{-dv.x, -dv.y, -dv.z};
        }
        best = fmax(best, 1.0 / accd_toc(Q, D, true, eta));
    });
    if (!ok) atomicOr(err, 1);
    inv_toc[e] = best;
    cnt[e] = m;
}

// ground (GIPC.cu:6945-7020): coef / (dist slackness) for vertices moving
// toward the plane
__global__ void k_ground_step(int n, const uint32_t* __restrict__ surf, const double* __restrict__ x,
                              const double* __restrict__ nrm, double offset, const double* __restrict__ dir,
                              double slackness, double* __restrict__ inv_t)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const V3 nv{nrm[0], nrm[1], nrm[2]};
    const int v = (int)surf[q];
    const double coef = dot(nv, ld3(dir, v));
    double t = 1.0;
    if (coef > 0.0) t = coef / ((dot(nv, ld3(x, v)) - offset) * slackness);
    inv_t[q] = t;
}

// tet injectivity (GIPC.cu:1130-1300,7022-7094): the smallest t > 0 at which
// the tet's volume shrinks to slackness times its current one; the
// reference's quadratic and linear branches below error_rate, the smallest
// positive root of the cubic otherwise (a case the reference leaves undefined)
__global__ void k_injective_step(int n, const uint32_t* __restrict__ tets, const double* __restrict__ x,
                                 const double* __restrict__ dir, double slackness, double error_rate,
                                 double* __restrict__ inv_t)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    int t[4];
    for (int a = 0; a < 4; ++a) t[a] = (int)tets[4L * e + a];
    const V3 x0 = ld3(x, t[0]), d0 = ld3(dir, t[0]);
    V3 A[3], Bm[3];
    for (int c = 0; c < 3; ++c) {
        A[c] = ld3(x, t[c + 1]) - x0;
        const V3 dc = ld3(dir, t[c + 1]);
        Bm[c] = V3{-dc.x + d0.x, -dc.y + d0.y, -dc.z + d0.z};
    }
    const double c0 = det_cols(A[0], A[1], A[2]), c3 = det_cols(Bm[0], Bm[1], Bm[2]);
    const double c1 = det_cols(Bm[0], A[1], A[2]) + det_cols(A[0], Bm[1], A[2]) + det_cols(A[0], A[1], Bm[2]);
    const double c2 = det_cols(A[0], Bm[1], Bm[2]) + det_cols(Bm[0], A[1], Bm[2]) + det_cols(Bm[0], Bm[1], A[2]);
    const double k0 = (1.0 - slackness) * c0;
    double ts = 1.0;
    if (fabs(c3) <= error_rate) {
        if (fabs(c2) <= error_rate) {
            ts = -k0 / c1;
        } else {
            const double disc = c1 * c1 - 4.0 * c2 * k0;
            if (disc > 0) {
                ts = (-c1 - sqrt(disc)) / (2.0 * c2);
                if (ts < 0) ts = (-c1 + sqrt(disc)) / (2.0 * c2);
            }
        }
    } else {
        double best = -1.0, tlo = 0.0, glo = k0;
        for (int s = 1; s <= 256 && best < 0.0; ++s) {
            const double th = 4.0 * s / 256;
            const double gh = ((c3 * th + c2) * th + c1) * th + k0;
            if ((glo > 0) != (gh > 0) || gh == 0.0) {
                double lo = tlo, hi = th;
                for (int it = 0; it < 100; ++it) {
                    const double mm = 0.5 * (lo + hi), gm = ((c3 * mm + c2) * mm + c1) * mm + k0;
                    if ((gm > 0) == (glo > 0)) lo = mm;
                    else hi = mm;
                }
                best = 0.5 * (lo + hi);
            }
            tlo = th;
            glo = gh;
        }
        ts = best > 0.0 ? best : 1.0;
    }
    if (ts <= 0.0) ts = 1.0;
    inv_t[e] = 1.0 / ts;
}

// max of n values and sum of n ints -> out[blockIdx.x], icnt[blockIdx.x]
__global__ __launch_bounds__(256) void k_max_reduce(long n, const double* __restrict__ v, const int* __restrict__ c,
                                                    double* __restrict__ out, int* __restrict__ icnt)
{
    __shared__ double sm[4];
    __shared__ int si[4];
    double m = 1.0;
    int k = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        m = fmax(m, v[i]);
        if (c != nullptr) k += c[i];
    }
    for (int off = 32; off >= 1; off >>= 1) {
        m = fmax(m, __shfl_xor(m, off, 64));
        k += __shfl_xor(k, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sm[threadIdx.x >> 6] = m;
        si[threadIdx.x >> 6] = k;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sm[0];
        int rk = si[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            r = fmax(r, sm[w]);
            rk += si[w];
        }
        out[blockIdx.x] = r;
        icnt[blockIdx.x] = rk;
    }
}

}  // namespace

void Lbvh::build(hipStream_t s, const double* x, const uint32_t* prims, int n_prims, int k, const double* dir,
                 double alpha)
{
    n = n_prims;
    if (n <= 0) return;
    const long nodes = 2L * n - 1;
    box.reserve(6 * nodes);
    leaf_box.reserve(6L * n);
    const int gb = std::min(g256(n), 256);
    scene.reserve(6L * (gb + 1));
    keys.reserve(n);
    skeys.reserve(n);
    idx.reserve(n);
    sidx.reserve(n);
    child.reserve(2L * (n > 1 ? n - 1 : 1));
    parent.reserve(nodes);
    flag.reserve(n > 1 ? n - 1 : 1);
    k_prim_boxes<<<g256(n), 256, 0, s>>>(n, x, prims, k, dir, alpha, leaf_box.get());
    SGIPC_LAUNCH_CHECK();
    k_box_union<<<gb, 256, 0, s>>>(n, leaf_box.get(), scene.get() + 6);
    SGIPC_LAUNCH_CHECK();
    k_box_union<<<1, 256, 0, s>>>(gb, scene.get() + 6, scene.get());
    SGIPC_LAUNCH_CHECK();
    k_morton<<<g256(n), 256, 0, s>>>(n, leaf_box.get(), scene.get(), keys.get(), idx.get());
    SGIPC_LAUNCH_CHECK();
    // the index in the low word keeps keys unique; a stable sort on the
    // Morton bits alone orders them the same way
    size_t need = 0;
    SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, need, keys.get(), skeys.get(), idx.get(), sidx.get(), (size_t)n, 32,
                                        62, s));
    temp.reserve(need);
    SGIPC_HIP(rocprim::radix_sort_pairs(temp.get(), need, keys.get(), skeys.get(), idx.get(), sidx.get(), (size_t)n,
                                        32, 62, s));
    launch_fill<int>(s, parent.get(), -1, nodes);
    if (n > 1) {
        launch_fill<int>(s, flag.get(), 0, n - 1);
        k_radix_tree<<<g256(n - 1), 256, 0, s>>>(n, skeys.get(), child.get(), parent.get());
        SGIPC_LAUNCH_CHECK();
    }
    k_refit<<<g256(n), 256, 0, s>>>(n, sidx.get(), leaf_box.get(), box.get(), child.get(), parent.get(), flag.get());
    SGIPC_LAUNCH_CHECK();
}

ContactCounts ContactDetector::detect(hipStream_t s, const double* x, const double* rest, int n_verts,
                                      const uint32_t* faces, int n_faces, const uint32_t* edges, int n_edges,
                                      const uint32_t* surf, int n_surf, const int* body_id, const int* btype,
                                      double dhat_sq, int* pairs, int* mat_index, int capacity)
{
    (void)n_verts;
    ContactCounts out;
    last_total = 0;
    const int nq_vf = (n_faces > 0) ? n_surf : 0;
    const int nq_ee = n_edges > 1 ? n_edges : 0;
    const long nq = (long)nq_vf + nq_ee;
    if (nq == 0) return out;
    if (nq_vf > 0) fbvh_.build(s, x, faces, n_faces, 3);
    if (nq_ee > 0) ebvh_.build(s, x, edges, n_edges, 2);
    cnt_.reserve(nq + 1);
    off_.reserve(nq + 1);
    err_.reserve(1);
    launch_fill<int>(s, err_.get(), 0, 1);
    launch_fill<int>(s, cnt_.get() + nq, 0, 1);
    if (nq_vf > 0) {
        k_query_vf<false><<<g256(nq_vf), 256, 0, s>>>(nq_vf, surf, x, faces, n_faces, fbvh_.box.get(),
                                                      fbvh_.child.get(), fbvh_.sidx.get(), body_id, btype, dhat_sq,
                                                      cnt_.get(), nullptr, nullptr, nullptr, nullptr, err_.get());
        SGIPC_LAUNCH_CHECK();
    }
    if (nq_ee > 0) {
        k_query_ee<false><<<g256(nq_ee), 256, 0, s>>>(nq_ee, edges, x, rest, ebvh_.box.get(), ebvh_.child.get(),
                                                      ebvh_.sidx.get(), ebvh_.leaf_box.get(), body_id, btype, dhat_sq,
                                                      cnt_.get() + nq_vf, nullptr, nullptr, nullptr, nullptr,
                                                      err_.get());
        SGIPC_LAUNCH_CHECK();
    }
    size_t need = 0;
    SGIPC_HIP(rocprim::exclusive_scan(nullptr, need, cnt_.get(), off_.get(), 0, (size_t)(nq + 1), rocprim::plus<int>(),
                                      s));
    temp_.reserve(need);
    SGIPC_HIP(rocprim::exclusive_scan(temp_.get(), need, cnt_.get(), off_.get(), 0, (size_t)(nq + 1),
                                      rocprim::plus<int>(), s));
    int h[3] = {0, 0, 0};
    SGIPC_HIP(hipMemcpyAsync(&h[0], off_.get() + nq_vf, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipMemcpyAsync(&h[1], off_.get() + nq, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipMemcpyAsync(&h[2], err_.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    SGIPC_CHECK(h[2] == 0, SGIPC_CAPACITY, "BVH traversal deeper than its stack");
    const int n_vf = h[0], total = h[1];
    out.c[0] = total;
    last_total = total;
    SGIPC_CHECK(total <= capacity, SGIPC_CAPACITY, "contact pairs exceed the output capacity");
    if (total == 0) return out;
    rkey_.reserve(total);
    rkey_sorted_.reserve(total);
    rid_.reserve(total);
    rid_sorted_.reserve(total);
    rpair_.reserve(total);
    rcls_.reserve(total);
    if (nq_vf > 0) {
        k_query_vf<true><<<g256(nq_vf), 256, 0, s>>>(nq_vf, surf, x, faces, n_faces, fbvh_.box.get(),
                                                     fbvh_.child.get(), fbvh_.sidx.get(), body_id, btype, dhat_sq,
                                                     nullptr, off_.get(), rkey_.get(), rpair_.get(), rcls_.get(),
                                                     err_.get());
        SGIPC_LAUNCH_CHECK();
    }
    if (nq_ee > 0) {
        k_query_ee<true><<<g256(nq_ee), 256, 0, s>>>(nq_ee, edges, x, rest, ebvh_.box.get(), ebvh_.child.get(),
                                                     ebvh_.sidx.get(), ebvh_.leaf_box.get(), body_id, btype, dhat_sq,
                                                     nullptr, off_.get() + nq_vf, rkey_.get(), rpair_.get(),
                                                     rcls_.get(), err_.get());
        SGIPC_LAUNCH_CHECK();
    }
    // canonical order: each query's segment sorted by its key
    k_iota<<<g256(total), 256, 0, s>>>(total, rid_.get());
    SGIPC_LAUNCH_CHECK();
    const int seg[2][2] = {{0, n_vf}, {n_vf, total}};
    for (const auto& sg : seg) {
        const int m = sg[1] - sg[0];
        if (m <= 0) continue;
        size_t nd = 0;
        SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, nd, rkey_.get() + sg[0], rkey_sorted_.get() + sg[0],
                                            rid_.get() + sg[0], rid_sorted_.get() + sg[0], (size_t)m, 0, 64, s));
        temp_.reserve(nd);
        SGIPC_HIP(rocprim::radix_sort_pairs(temp_.get(), nd, rkey_.get() + sg[0], rkey_sorted_.get() + sg[0],
                                            rid_.get() + sg[0], rid_sorted_.get() + sg[0], (size_t)m, 0, 64, s));
    }
    flags_.reserve(3L * total);
    scan_.reserve(3L * total);
    k_emit<<<g256(total), 256, 0, s>>>(total, rid_sorted_.get(), rpair_.get(), rcls_.get(),
                                       reinterpret_cast<int4*>(pairs), flags_.get());
    SGIPC_LAUNCH_CHECK();
    int cls_tot[3] = {0, 0, 0};
    for (int c = 0; c < 3; ++c) {
        size_t nd = 0;
        SGIPC_HIP(rocprim::exclusive_scan(nullptr, nd, flags_.get() + (long)c * total, scan_.get() + (long)c * total, 0,
                                          (size_t)total, rocprim::plus<int>(), s));
        temp_.reserve(nd);
        SGIPC_HIP(rocprim::exclusive_scan(temp_.get(), nd, flags_.get() + (long)c * total,
                                          scan_.get() + (long)c * total, 0, (size_t)total, rocprim::plus<int>(), s));
    }
    k_mat_index<<<g256(total), 256, 0, s>>>(total, rcls_.get(), rid_sorted_.get(), scan_.get(), mat_index);
    SGIPC_LAUNCH_CHECK();
    int last[6];
    for (int c = 0; c < 3; ++c) {
        SGIPC_HIP(hipMemcpyAsync(&last[2 * c], scan_.get() + (long)c * total + total - 1, sizeof(int),
                                 hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipMemcpyAsync(&last[2 * c + 1], flags_.get() + (long)c * total + total - 1, sizeof(int),
                                 hipMemcpyDeviceToHost, s));
    }
    SGIPC_HIP(hipStreamSynchronize(s));
    for (int c = 0; c < 3; ++c) cls_tot[c] = last[2 * c] + last[2 * c + 1];
    out.c[2] = cls_tot[0];
    out.c[3] = cls_tot[1];
    out.c[4] = cls_tot[2];
    return out;
}


double StepReducer::max_inv(hipStream_t s, const double* v, const int* c, long n, long* total)
{
    const int gb = (int)std::min<long>(256, (n + 255) / 256);
    part_.reserve(gb + 1);
    ipart_.reserve(gb + 1);
    k_max_reduce<<<gb, 256, 0, s>>>(n, v, c, part_.get(), ipart_.get());
    SGIPC_LAUNCH_CHECK();
    k_max_reduce<<<1, 256, 0, s>>>(gb, part_.get(), ipart_.get(), part_.get() + gb, ipart_.get() + gb);
    SGIPC_LAUNCH_CHECK();
    double h = 1.0;
    int hk = 0;
    SGIPC_HIP(hipMemcpyAsync(&h, part_.get() + gb, sizeof(double), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipMemcpyAsync(&hk, ipart_.get() + gb, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    if (total) *total = hk;
    return h;
}

double ContactDetector::ccd_step(hipStream_t s, const double* x, const double* dir, double alpha,
                                 const uint32_t* faces, int n_faces, const uint32_t* edges, int n_edges,
                                 const uint32_t* surf, int n_surf, const int* body_id, const int* btype,
                                 double dhat_sq, double slackness, long* n_candidates)
{
    const double eta = 1.0 - slackness;
    const int nq_vf = n_faces > 0 ? n_surf : 0;
    const int nq_ee = n_edges > 1 ? n_edges : 0;
    double inv = 1.0;
    long cand = 0;
    err_.reserve(1);
    launch_fill<int>(s, err_.get(), 0, 1);
    if (nq_vf > 0) {
        fbvh_.build(s, x, faces, n_faces, 3, dir, alpha);
        vals_.reserve(nq_vf);
        cnt_.reserve(nq_vf);
        k_ccd_vf<<<g256(nq_vf), 256, 0, s>>>(nq_vf, surf, x, dir, alpha, faces, n_faces, fbvh_.box.get(),
                                             fbvh_.child.get(), fbvh_.sidx.get(), body_id, btype, dhat_sq, eta,
                                             vals_.get(), cnt_.get(), err_.get());
        SGIPC_LAUNCH_CHECK();
        long c = 0;
        inv = std::max(inv, red_.max_inv(s, vals_.get(), cnt_.get(), nq_vf, &c));
        cand += c;
    }
    if (nq_ee > 0) {
        ebvh_.build(s, x, edges, n_edges, 2, dir, alpha);
        vals_.reserve(nq_ee);
        cnt_.reserve(nq_ee);
        k_ccd_ee<<<g256(nq_ee), 256, 0, s>>>(nq_ee, edges, x, dir, ebvh_.box.get(), ebvh_.child.get(),
                                             ebvh_.sidx.get(), ebvh_.leaf_box.get(), body_id, btype, dhat_sq, eta,
                                             vals_.get(), cnt_.get(), err_.get());
        SGIPC_LAUNCH_CHECK();
        long c = 0;
        inv = std::max(inv, red_.max_inv(s, vals_.get(), cnt_.get(), nq_ee, &c));
        cand += c;
    }
    int h = 0;
    SGIPC_HIP(hipMemcpyAsync(&h, err_.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    SGIPC_CHECK(h == 0, SGIPC_CAPACITY, "BVH traversal deeper than its stack");
    if (n_candidates) *n_candidates = cand;
    return 1.0 / inv;
}

double ground_step(hipStream_t s, StepReducer& red, DevBuf<double>& vals, const double* x, const uint32_t* surf,
                   int n_surf, const double* normal, double offset, const double* dir, double slackness)
{
    if (n_surf <= 0) return 1.0;
    vals.reserve(n_surf);
    k_ground_step<<<g256(n_surf), 256, 0, s>>>(n_surf, surf, x, normal, offset, dir, slackness, vals.get());
    SGIPC_LAUNCH_CHECK();
    return 1.0 / red.max_inv(s, vals.get(), nullptr, n_surf, nullptr);
}

double injective_step(hipStream_t s, StepReducer& red, DevBuf<double>& vals, const double* x, const uint32_t* tets,
                      int n_tets, const double* dir, double slackness, double error_rate)
{
    if (n_tets <= 0) return 1.0;
    vals.reserve(n_tets);
    k_injective_step<<<g256(n_tets), 256, 0, s>>>(n_tets, tets, x, dir, slackness, error_rate, vals.get());
    SGIPC_LAUNCH_CHECK();
    return 1.0 / red.max_inv(s, vals.get(), nullptr, n_tets, nullptr);
}

void launch_barrier_pairs(hipStream_t s, int n_pairs, const double* x, const int* pairs, const int* mat_index, int n4,
                          int n3, double dhat_sq, double kappa, int hess_offset, double* grad, double* tvals,
                          int* trows, int* tcols, int* bad)
{
    if (n_pairs <= 0) return;
    k_barrier_pairs<<<(n_pairs + 127) / 128, 128, 0, s>>>(n_pairs, x, reinterpret_cast<const int4*>(pairs), mat_index,
                                                          n4, n3, dhat_sq, kappa, hess_offset, grad, tvals, trows,
                                                          tcols, bad);
    SGIPC_LAUNCH_CHECK();
}

}  // namespace sgipc
