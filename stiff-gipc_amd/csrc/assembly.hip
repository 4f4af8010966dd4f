// Hessian assembly for gfx950: element kernels that write 3x3 triplets and
// the triplet -> unique upper BCOO conversion.
//
//  * stable Neo-Hookean tet (StiffGIPC/femEnergy.cu:1719-1857, PK1 :432-459,
//    projection :1641-1717): the PSD projection uses the analytic eigensystem
//    of dP/dF (three twist, three flip and three scaling modes built from a
//    rotation-variant 3x3 SVD), accumulated directly in the 12x12 vertex
//    space as sum_i max(l_i,0) g_i g_i^T with g_i = dF/dx^T vec(Q_i), so no
//    9x9 or 12x12 dense product is ever formed;
//  * point-triangle barrier, GIPC rank-2 variant (StiffGIPC/GIPC.cu:4733-4984):
//    rank-1 Hessian lambda0(I5) g g^T with g the gradient of dist/sqrt(dHat);
//    the lower blocks are moved to (col,row) and zeroed as the reference's
//    contact partition does (StiffGIPC/GIPC.cu:10329-10363);
//  * mass diagonal (StiffGIPC/GIPC.cu:10466-10485);
//  * conversion (Converter::srbk_convert, StiffGIPC/linear_system/utils/
//    converter.cu:32-238): a stable rocPRIM radix sort over only the bits the
//    (row, col) key needs, then every unique block is summed by one thread in
//    sorted (= arena) order, a block of more than 64 triplets as the ordered
//    sum of its 64-triplet chunk sums (each chunk by its own thread), which
//    makes the result deterministic.
#include <algorithm>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "assembly.hpp"
#include "pcg.hpp"
#include "wave_ops.hpp"

namespace sgipc {

namespace {

// ------------------------------------------------------------------ 3x3 helpers
__device__ __forceinline__ void jacobi3(double A[3][3], double V[3][3])
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 10; ++sweep) {
        const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        const double dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off <= 1e-34 * dia || off == 0.0) break;
#pragma unroll
        for (int pq = 0; pq < 3; ++pq) {
            const int p = pq == 2 ? 1 : 0;
            const int q = pq == 0 ? 1 : 2;
            const double apq = A[p][q];
            if (apq == 0.0) continue;
            const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
            const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double akp = A[k][p], akq = A[k][q];
                A[k][p] = c * akp - s * akq;
                A[k][q] = s * akp + c * akq;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double apk = A[p][k], aqk = A[q][k];
                A[p][k] = c * apk - s * aqk;
                A[q][k] = s * apk + c * aqk;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double vkp = V[k][p], vkq = V[k][q];
                V[k][p] = c * vkp - s * vkq;
                V[k][q] = s * vkp + c * vkq;
            }
        }
    }
}

__device__ __forceinline__ double det3(const double F[3][3])
{
    return F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) - F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
           F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
}

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double c[3])
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// Rotation-variant SVD F = U diag(S) V^T with det U = det V = +1 (S[2] may be
// negative), from the eigen-decomposition of F^T F.  Columns of U/V are u_i/v_i.
__device__ void svd3_rv(const double F[3][3], double U[3][3], double S[3], double V[3][3])
{
    double B[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) B[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
    jacobi3(B, V);
    double ev[3] = {B[0][0], B[1][1], B[2][2]};
    // sort eigenvalues descending (swap V columns alongside)
#pragma unroll
    for (int pass = 0; pass < 2; ++pass)
#pragma unroll
        for (int i = 0; i < 2; ++i)
            if (ev[i] < ev[i + 1]) {
                double t = ev[i]; ev[i] = ev[i + 1]; ev[i + 1] = t;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    double w = V[k][i]; V[k][i] = V[k][i + 1]; V[k][i + 1] = w;
                }
            }
    if (det3(V) < 0.0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) V[k][2] = -V[k][2];
    }
    double fv[3][3];  // F v_i as columns
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int r = 0; r < 3; ++r) fv[r][i] = F[r][0] * V[0][i] + F[r][1] * V[1][i] + F[r][2] * V[2][i];
    double u0[3] = {fv[0][0], fv[1][0], fv[2][0]};
    double n0 = sqrt(u0[0] * u0[0] + u0[1] * u0[1] + u0[2] * u0[2]);
    if (n0 > 1e-300) {
#pragma unroll
        for (int k = 0; k < 3; ++k) u0[k] /= n0;
    } else {
        u0[0] = 1.0; u0[1] = 0.0; u0[2] = 0.0;
    }
    double u1[3] = {fv[0][1], fv[1][1], fv[2][1]};
    const double pr = u1[0] * u0[0] + u1[1] * u0[1] + u1[2] * u0[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) u1[k] -= pr * u0[k];
    double n1 = sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    if (n1 > 1e-14 * (n0 > 0 ? n0 : 1.0)) {
#pragma unroll
        for (int k = 0; k < 3; ++k) u1[k] /= n1;
    } else {
        // any unit vector orthogonal to u0
        double e[3] = {fabs(u0[0]) < 0.9 ? 1.0 : 0.0, fabs(u0[0]) < 0.9 ? 0.0 : 1.0, 0.0};
        cross3(u0, e, u1);
        double nn = sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
#pragma unroll
        for (int k = 0; k < 3; ++k) u1[k] /= nn;
    }
    double u2[3];
    cross3(u0, u1, u2);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        U[k][0] = u0[k];
        U[k][1] = u1[k];
        U[k][2] = u2[k];
    }
    S[0] = u0[0] * fv[0][0] + u0[1] * fv[1][0] + u0[2] * fv[2][0];
    S[1] = u1[0] * fv[0][1] + u1[1] * fv[1][1] + u1[2] * fv[2][1];
    S[2] = u2[0] * fv[0][2] + u2[1] * fv[1][2] + u2[2] * fv[2][2];
}

__device__ __forceinline__ int up12(int r, int c) { return r * 12 - r * (r - 1) / 2 + (c - r); }

// acc += lam * g g^T (upper triangle of 12x12)
__device__ __forceinline__ void rank1_upper(double acc[78], const double g[12], double lam)
{
#pragma unroll
    for (int r = 0; r < 12; ++r) {
        const double lr = lam * g[r];
#pragma unroll
        for (int c = r; c < 12; ++c) acc[up12(r, c)] += lr * g[c];
    }
}

// ------------------------------------------------------------------ tet kernel
__global__ __launch_bounds__(128) void k_tet_snh(int n_tets, const double* __restrict__ x,
                                                 const uint4* __restrict__ tets, const double* __restrict__ dm_inv,
                                                 const double* __restrict__ vol, const double* __restrict__ mu_,
                                                 const double* __restrict__ lam_, double dt2, int hess_offset,
                                                 double* __restrict__ grad, double* __restrict__ tvals,
                                                 int* __restrict__ trows, int* __restrict__ tcols)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tets) return;
    const uint4 tv = tets[t];
    const unsigned vid[4] = {tv.x, tv.y, tv.z, tv.w};
    double xs[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) xs[a][d] = x[3 * (long)vid[a] + d];
    double Dmi[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Dmi[i][j] = dm_inv[9 * (long)t + 3 * i + j];
    // F = Ds Dm^-1, Ds columns x1-x0, x2-x0, x3-x0
    double F[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) s += (xs[k + 1][i] - xs[0][i]) * Dmi[k][j];
            F[i][j] = s;
        }
    // d_a: dF(:,j)/dx_a coefficients
    double D[4][3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        D[0][j] = -(Dmi[0][j] + Dmi[1][j] + Dmi[2][j]);
#pragma unroll
        for (int a = 1; a < 4; ++a) D[a][j] = Dmi[a - 1][j];
    }
    const double mu = mu_[t], lam = lam_[t];
    const double J = det3(F);
    const double c = lam * (J - 1.0) - mu;
    const double scale = dt2 * vol[t];

    if (grad) {
        double cof[3][3];
        cof[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
        cof[0][1] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
        cof[0][2] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
        cof[1][0] = F[2][1] * F[0][2] - F[2][2] * F[0][1];
        cof[1][1] = F[2][2] * F[0][0] - F[2][0] * F[0][2];
        cof[1][2] = F[2][0] * F[0][1] - F[2][1] * F[0][0];
        cof[2][0] = F[0][1] * F[1][2] - F[1][1] * F[0][2];
        cof[2][1] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
        cof[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < 3; ++j) s += (mu * F[d][j] + c * cof[d][j]) * D[a][j];
                atomic_add(grad + 3 * (long)vid[a] + d, scale * s);
            }
    }

    double U[3][3], S[3], V[3][3];
    svd3_rv(F, U, S, V);
    // w[a][i] = v_i . d_a
    double w[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i) w[a][i] = V[0][i] * D[a][0] + V[1][i] * D[a][1] + V[2][i] * D[a][2];

    double acc[78];
#pragma unroll
    for (int e = 0; e < 78; ++e) acc[e] = 0.0;
    const double rs2 = 0.70710678118654752440;
    // twist (Q = (u_j v_i^T - u_i v_j^T)/sqrt2) and flip (Q = (u_i v_j^T + u_j v_i^T)/sqrt2)
    // modes for axis k with (i, j) the other two axes; eigenvalues mu +/- c S[k]
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int i = k == 0 ? 1 : 0;
        const int j = k == 2 ? 1 : 2;
#pragma unroll
        for (int flip = 0; flip < 2; ++flip) {
            const double l = flip ? (mu - c * S[k]) : (mu + c * S[k]);
            if (l <= 0.0) continue;
            double g[12];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    g[3 * a + d] = flip ? rs2 * (U[d][i] * w[a][j] + U[d][j] * w[a][i])
                                        : rs2 * (U[d][j] * w[a][i] - U[d][i] * w[a][j]);
            rank1_upper(acc, g, l);
        }
    }
    // scaling modes: eigen-decomposition of the 3x3 block in singular space
    {
        const double ev = lam * (2.0 * J - 1.0) - mu;
        double A[3][3];
        A[0][0] = mu + lam * S[1] * S[1] * S[2] * S[2];
        A[1][1] = mu + lam * S[0] * S[0] * S[2] * S[2];
        A[2][2] = mu + lam * S[0] * S[0] * S[1] * S[1];
        A[0][1] = A[1][0] = ev * S[2];
        A[0][2] = A[2][0] = ev * S[1];
        A[1][2] = A[2][1] = ev * S[0];
        double E[3][3];
        jacobi3(A, E);
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const double l = A[m][m];
            if (l <= 0.0) continue;
            double g[12];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int d = 0; d < 3; ++d)
                    g[3 * a + d] = E[0][m] * U[d][0] * w[a][0] + E[1][m] * U[d][1] * w[a][1] + E[2][m] * U[d][2] * w[a][2];
            rank1_upper(acc, g, l);
        }
    }
    // 10 upper blocks: local (ii <= jj); stored at (min,max) of the global ids
    const long base = 10 * (long)t;
    int kk = 0;
#pragma unroll
    for (int ii = 0; ii < 4; ++ii)
#pragma unroll
        for (int jj = ii; jj < 4; ++jj, ++kk) {
            const int row = (int)vid[ii] + hess_offset, col = (int)vid[jj] + hess_offset;
            double* dst = tvals + 9 * (base + kk);
            const bool keep = row <= col;
            trows[base + kk] = keep ? row : col;
            tcols[base + kk] = keep ? col : row;
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) {
                    // H(3ii+r, 3jj+cc) from the upper accumulator
                    int R = 3 * ii + r, C = 3 * jj + cc;
                    double hv = (R <= C) ? acc[up12(R, C)] : acc[up12(C, R)];
                    // keep: M(r,cc) = H(R,C); else M(r,cc) = H(3jj+r, 3ii+cc) = H(C',R') transposed
                    int R2 = 3 * jj + r, C2 = 3 * ii + cc;
                    double ht = (R2 <= C2) ? acc[up12(R2, C2)] : acc[up12(C2, R2)];
                    dst[cc * 3 + r] = scale * (keep ? hv : ht);
                }
        }
}

// ------------------------------------------------------------------ barrier kernel
__device__ __forceinline__ double barrier_lambda0(double I5, double dh, double kappa)
{
    const double L = log(I5);
    return -(4.0 * kappa * dh * dh *
             (4.0 * I5 + L - 3.0 * I5 * I5 * L * L + 6.0 * I5 * L - 2.0 * I5 * I5 + I5 * L * L - 7.0 * I5 * I5 * L - 2.0)) /
           I5;
}

__global__ __launch_bounds__(128) void k_barrier_pt(int n_pairs, const double* __restrict__ x,
                                                    const int4* __restrict__ pairs, const double* __restrict__ pair_x,
                                                    double dh, double kappa, int hess_offset,
                                                    double* __restrict__ grad, double* __restrict__ tvals,
                                                    int* __restrict__ trows, int* __restrict__ tcols)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pairs) return;
    const int4 pr = pairs[p];
    const int vid[4] = {-pr.x - 1, pr.y, pr.z, pr.w};
    double P[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d)
            P[a][d] = pair_x ? pair_x[12 * (long)p + 3 * a + d] : x[3 * (long)vid[a] + d];
    double e1[3], e2[3], w[3], b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        e1[d] = P[2][d] - P[1][d];
        e2[d] = P[3][d] - P[1][d];
        w[d] = P[0][d] - P[1][d];
    }
    cross3(e1, e2, b);
    const double bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    const double aTb = w[0] * b[0] + w[1] * b[1] + w[2] * b[2];
    const double d2 = aTb * aTb / bb;
    const double dis = sqrt(d2);
    const double sq = sqrt(dh);
    const double I5 = d2 / dh;
    const double nb = sqrt(bb), sgn = aTb >= 0 ? 1.0 : -1.0;
    double c1[3], c2[3];
    cross3(w, e2, c1);
    cross3(e1, w, c2);
    const double be1 = (c1[0] * b[0] + c1[1] * b[1] + c1[2] * b[2]) / bb;
    const double be2 = (c2[0] * b[0] + c2[1] * b[1] + c2[2] * b[2]) / bb;
    const double be0 = 1.0 - be1 - be2;
    const double coef[4] = {1.0, -be0, -be1, -be2};
    double g[12];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) g[3 * a + d] = sgn * coef[a] * b[d] / (nb * sq);
    const double gass = 1e-6;
    const double lam0 = barrier_lambda0(d2 < gass * dh ? gass : I5, dh, kappa);
    if (grad) {
        const double L = log(I5);
        const double pk = 2.0 * (2.0 * kappa * dh * dh * L * (I5 - 1.0) * (I5 + I5 * L - 1.0)) / I5;
        const double s = dis / sq;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) atomic_add(grad + 3 * (long)vid[a] + d, g[3 * a + d] * s * pk);
    }
    const long base = 16 * (long)p;
#pragma unroll
    for (int ii = 0; ii < 4; ++ii)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            const int kk = ii * 4 + jj;
            const int row = vid[ii] + hess_offset, col = vid[jj] + hess_offset;
            double* dst = tvals + 9 * (base + kk);
            if (row <= col) {
                trows[base + kk] = row;
                tcols[base + kk] = col;
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) dst[cc * 3 + r] = lam0 * g[3 * ii + r] * g[3 * jj + cc];
            } else {
                trows[base + kk] = col;
                tcols[base + kk] = row;
#pragma unroll
                for (int e = 0; e < 9; ++e) dst[e] = 0.0;
            }
        }
}

__global__ void k_mass_diag(int n, const double* __restrict__ mass, int hess_offset, double* __restrict__ tvals,
                            int* __restrict__ trows, int* __restrict__ tcols)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = mass[i];
    double* dst = tvals + 9 * (long)i;
#pragma unroll
    for (int e = 0; e < 9; ++e) dst[e] = (e == 0 || e == 4 || e == 8) ? m : 0.0;
    trows[i] = i + hess_offset;
    tcols[i] = i + hess_offset;
}

// ------------------------------------------------------------------ conversion
__global__ void k_make_keys(const int* __restrict__ rows, const int* __restrict__ cols, long n, int kbits,
                            uint64_t* __restrict__ keys, int* __restrict__ idx, int* __restrict__ bad)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = rows[i], c = cols[i];
    if (r < 0 || c < 0 || r > c || (kbits < 31 && (r >> kbits) != 0) || (kbits < 31 && (c >> kbits) != 0)) atomicOr(bad, 1);
    keys[i] = ((uint64_t)(uint32_t)r << kbits) | (uint64_t)(uint32_t)c;
    idx[i] = (int)i;
}

__global__ void k_head_flags(const uint64_t* __restrict__ sk, long n, int* __restrict__ flag)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || sk[i] != sk[i - 1]) ? 1 : 0;
}

__global__ void k_heads(const uint64_t* __restrict__ sk, const int* __restrict__ flag, const int* __restrict__ pos,
                        long n, int kbits, int* __restrict__ head_pos, int* __restrict__ orows,
                        int* __restrict__ ocols, int capacity)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i == n - 1 && pos[i] <= capacity) head_pos[pos[i]] = (int)n;  // sentinel
    if (!flag[i]) return;
    const int u = pos[i] - 1;
    if (u > capacity) return;
    // u == capacity: the first block past the capacity, whose start ends the
    // last block kept (head_pos holds capacity + 1 entries)
    head_pos[u] = (int)i;
    if (u == capacity) return;
    const uint64_t k = sk[i];
    orows[u] = (int)(k >> kbits);
    ocols[u] = (int)(k & ((1ull << kbits) - 1ull));
}

// Unique blocks of more than kSegChunk triplets (on coarse MAS levels and
// affine-body rows thousands of triplets merge into one block) are summed in
// chunks of kSegChunk consecutive sorted triplets, each by its own thread,
// and then the chunk sums in order; shorter blocks by one thread in sorted
// order.  Fixed order either way (deterministic).
constexpr int kSegChunk = 64;

// chunks of every unique block u < cap (0: summed directly)
__global__ void k_seg_chunks(const int* __restrict__ head_pos, const int* __restrict__ pos, long n, int cap,
                             int* __restrict__ cnt)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= cap) return;
    const int m = pos[n - 1] < cap ? pos[n - 1] : cap;
    int c = 0;
    if (u < m) {
        const int len = head_pos[u + 1] - head_pos[u];
        if (len > kSegChunk) c = (len + kSegChunk - 1) / kSegChunk;
    }
    cnt[u] = c;
}

// chunk t of the long blocks (cofs: exclusive prefix of cnt over cap blocks)
__global__ __launch_bounds__(256) void k_chunk_sum(const double* __restrict__ tvals, const int* __restrict__ sidx,
                                                   const int* __restrict__ head_pos, const int* __restrict__ cnt,
                                                   const int* __restrict__ cofs, int cap, long bound,
                                                   double* __restrict__ part)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= bound) return;
    const long total = (long)cofs[cap - 1] + cnt[cap - 1];
    if (t >= total) return;
    // the block holding chunk t: the last u with cofs[u] <= t
    int lo = 0, hi = cap;  // first index with cofs > t lies in (lo, hi]
    while (lo + 1 < hi) {
        const int mid = (lo + hi) >> 1;
        if ((long)cofs[mid] <= t) lo = mid;
        else hi = mid;
    }
    const int u = lo;
    const int b = head_pos[u] + (int)(t - cofs[u]) * kSegChunk;
    const int e = min(b + kSegChunk, head_pos[u + 1]);
    double s[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) s[q] = 0.0;
    for (int p = b; p < e; ++p) {
        const double* v = tvals + 9 * (long)sidx[p];
#pragma unroll
        for (int q = 0; q < 9; ++q) s[q] += v[q];
    }
    double* o = part + 9 * t;
#pragma unroll
    for (int q = 0; q < 9; ++q) o[q] = s[q];
}

// One thread per unique block: its triplets in sorted (= arena) order, or
// its chunk sums in order.
__global__ __launch_bounds__(256) void k_segment_sum(const double* __restrict__ tvals, const int* __restrict__ sidx,
                                                     const int* __restrict__ head_pos, const int* __restrict__ pos,
                                                     long n, double* __restrict__ ovals, int capacity,
                                                     const int* __restrict__ cnt, const int* __restrict__ cofs,
                                                     const double* __restrict__ part)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    const int nnz = pos[n - 1];
    if (u >= nnz || u >= capacity) return;
    double s[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) s[q] = 0.0;
    if (cnt[u] > 0) {
        const long c0 = cofs[u], c1 = c0 + cnt[u];
        for (long c = c0; c < c1; ++c) {
            const double* v = part + 9 * c;
#pragma unroll
            for (int q = 0; q < 9; ++q) s[q] += v[q];
        }
    } else {
        const int b = head_pos[u], e = head_pos[u + 1];
        for (int p = b; p < e; ++p) {
            const double* v = tvals + 9 * (long)sidx[p];
#pragma unroll
            for (int q = 0; q < 9; ++q) s[q] += v[q];
        }
    }
    double* o = ovals + 9 * (long)u;
#pragma unroll
    for (int q = 0; q < 9; ++q) o[q] = s[q];
}

__global__ void k_max_index(const int* __restrict__ rows, const int* __restrict__ cols, long n, int* __restrict__ out)
{
    int m = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        int r = rows[i], c = cols[i];
        m = r > m ? r : m;
        m = c > m ? c : m;
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

inline int g256(long n) { return (int)((n + 255) / 256); }

}  // namespace

void launch_tet_snh(hipStream_t s, int n_tets, const double* x, const uint32_t* tets, const double* dm_inv,
                    const double* vol, const double* mu, const double* lam, double dt, int hess_offset,
                    double* grad, double* tvals, int* trows, int* tcols)
{
    if (n_tets <= 0) return;
    k_tet_snh<<<(n_tets + 127) / 128, 128, 0, s>>>(n_tets, x, reinterpret_cast<const uint4*>(tets), dm_inv, vol, mu,
                                                   lam, dt * dt, hess_offset, grad, tvals, trows, tcols);
    SGIPC_LAUNCH_CHECK();
}

void launch_barrier_pt(hipStream_t s, int n_pairs, const double* x, const int* pairs, const double* pair_x,
                       double dhat_sq, double kappa, int hess_offset, double* grad, double* tvals, int* trows,
                       int* tcols)
{
    if (n_pairs <= 0) return;
    k_barrier_pt<<<(n_pairs + 127) / 128, 128, 0, s>>>(n_pairs, x, reinterpret_cast<const int4*>(pairs), pair_x,
                                                       dhat_sq, kappa, hess_offset, grad, tvals, trows, tcols);
    SGIPC_LAUNCH_CHECK();
}

void launch_mass_diag(hipStream_t s, int n, const double* mass, int hess_offset, double* tvals, int* trows,
                      int* tcols)
{
    if (n <= 0) return;
    k_mass_diag<<<g256(n), 256, 0, s>>>(n, mass, hess_offset, tvals, trows, tcols);
    SGIPC_LAUNCH_CHECK();
}

int convert_max_index(hipStream_t s, const int* rows, const int* cols, long n, int* ints)
{
    launch_fill<int>(s, ints, 0, 1);
    k_max_index<<<256, 256, 0, s>>>(rows, cols, n, ints);
    SGIPC_LAUNCH_CHECK();
    int m = 0;
    SGIPC_HIP(hipMemcpyAsync(&m, ints, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    return m;
}

size_t convert_temp_bytes(hipStream_t s, long n)
{
    size_t need = 0, need2 = 0;
    const int kbits = 31;
    SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, need, static_cast<uint64_t*>(nullptr), static_cast<uint64_t*>(nullptr),
                                        static_cast<int*>(nullptr), static_cast<int*>(nullptr), (size_t)n, 0, 2 * kbits,
                                        s));
    SGIPC_HIP(rocprim::inclusive_scan(nullptr, need2, static_cast<int*>(nullptr), static_cast<int*>(nullptr), (size_t)n,
                                      rocprim::plus<int>(), s));
    size_t need3 = 0;
    SGIPC_HIP(rocprim::exclusive_scan(nullptr, need3, static_cast<int*>(nullptr), static_cast<int*>(nullptr), 0,
                                      (size_t)n, rocprim::plus<int>(), s));
    return std::max(need, std::max(need2, need3));
}

int convert_into(hipStream_t s, const ConvertScratch& w, const double* tvals, const int* trows, const int* tcols,
                 long n, int index_bound, double* ovals, int* orows, int* ocols, int capacity, int* unique)
{
    *unique = 0;
    if (n <= 0) return 0;
    SGIPC_CHECK(n < (1L << 31), SGIPC_CAPACITY, "too many triplets");
    const int kbits = Converter::index_bits(index_bound);
    // at most n unique blocks: the kernels see the capacity clamped to n
    // (same results), so the block starts need n + 1 entries at most
    const int cap = capacity < n ? capacity : (int)n;
    launch_fill<int>(s, w.ints, 0, 1);
    k_make_keys<<<g256(n), 256, 0, s>>>(trows, tcols, n, kbits, w.keys, w.idx, w.ints);
    SGIPC_LAUNCH_CHECK();
    size_t need = 0, need2 = 0;
    SGIPC_HIP(rocprim::radix_sort_pairs(nullptr, need, w.keys, w.skeys, w.idx, w.sidx, (size_t)n, 0, 2 * kbits, s));
    SGIPC_HIP(rocprim::inclusive_scan(nullptr, need2, w.flag, w.pos, (size_t)n, rocprim::plus<int>(), s));
    SGIPC_CHECK(need <= w.temp_bytes && need2 <= w.temp_bytes, SGIPC_CAPACITY,
                "conversion scratch smaller than the sort needs");
    SGIPC_HIP(rocprim::radix_sort_pairs(w.temp, need, w.keys, w.skeys, w.idx, w.sidx, (size_t)n, 0, 2 * kbits, s));
    k_head_flags<<<g256(n), 256, 0, s>>>(w.skeys, n, w.flag);
    SGIPC_LAUNCH_CHECK();
    SGIPC_HIP(rocprim::inclusive_scan(w.temp, need2, w.flag, w.pos, (size_t)n, rocprim::plus<int>(), s));
    k_heads<<<g256(n), 256, 0, s>>>(w.skeys, w.flag, w.pos, n, kbits, w.head, orows, ocols, cap);
    SGIPC_LAUNCH_CHECK();
    if (cap > 0) {
        // long blocks in chunks: counts into flag, their exclusive prefix into
        // idx, the chunk sums into keys (both free after the sort and scan;
        // at most 2 n / kSegChunk chunks of 9 doubles fit in n words)
        int* cnt = w.flag;
        int* cofs = w.idx;
        double* part = reinterpret_cast<double*>(w.keys);
        k_seg_chunks<<<g256(cap), 256, 0, s>>>(w.head, w.pos, n, cap, cnt);
        SGIPC_LAUNCH_CHECK();
        size_t need3 = 0;
        SGIPC_HIP(rocprim::exclusive_scan(nullptr, need3, cnt, cofs, 0, (size_t)cap, rocprim::plus<int>(), s));
        SGIPC_CHECK(need3 <= w.temp_bytes, SGIPC_CAPACITY, "conversion scratch smaller than the scan needs");
        SGIPC_HIP(rocprim::exclusive_scan(w.temp, need3, cnt, cofs, 0, (size_t)cap, rocprim::plus<int>(), s));
        const long bound = 2 * n / kSegChunk + 1;
        k_chunk_sum<<<g256(bound), 256, 0, s>>>(tvals, w.sidx, w.head, cnt, cofs, cap, bound, part);
        SGIPC_LAUNCH_CHECK();
        k_segment_sum<<<g256(cap), 256, 0, s>>>(tvals, w.sidx, w.head, w.pos, n, ovals, cap, cnt, cofs, part);
        SGIPC_LAUNCH_CHECK();
    }
    int h[2] = {0, 0};
    SGIPC_HIP(hipMemcpyAsync(&h[0], w.pos + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipMemcpyAsync(&h[1], w.ints, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    SGIPC_CHECK(h[1] == 0, SGIPC_BAD_INDEX, "triplet with row > col, a negative index or an index beyond the bound");
    *unique = h[0];
    SGIPC_CHECK(h[0] <= capacity, SGIPC_CAPACITY, "unique block count exceeds the output capacity");
    return h[0];
}

int Converter::index_bits(int bound)
{
    int k = 1;
    while (k < 31 && (1L << k) < (long)bound) ++k;
    return k;
}

int Converter::max_index(hipStream_t s, const int* rows, const int* cols, long n)
{
    scratch_int_.reserve(4);
    return convert_max_index(s, rows, cols, n, scratch_int_.get());
}

int Converter::convert(hipStream_t s, const double* tvals, const int* trows, const int* tcols, long n,
                       int index_bound, double* ovals, int* orows, int* ocols, int capacity)
{
    if (n <= 0) return 0;
    SGIPC_CHECK(n < (1L << 31), SGIPC_CAPACITY, "too many triplets");
    keys_.reserve(n);
    skeys_.reserve(n);
    idx_.reserve(n);
    sidx_.reserve(n);
    flag_.reserve(n);
    pos_.reserve(n);
    head_.reserve((size_t)(capacity < n ? capacity : n) + 1);
    scratch_int_.reserve(4);
    temp_.reserve(convert_temp_bytes(s, n));
    ConvertScratch w;
    w.keys = keys_.get();
    w.skeys = skeys_.get();
    w.idx = idx_.get();
    w.sidx = sidx_.get();
    w.flag = flag_.get();
    w.pos = pos_.get();
    w.head = head_.get();
    w.ints = scratch_int_.get();
    w.temp = temp_.get();
    w.temp_bytes = temp_.n;
    return convert_into(s, w, tvals, trows, tcols, n, index_bound, ovals, orows, ocols, capacity, &last_unique);
}

}  // namespace sgipc
