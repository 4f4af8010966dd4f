// Lagged IPC friction of the contact pairs and of the ground, for gfx950.
//
// Reference: _calFrictionLastH_DistAndTan / _calFrictionLastH_gd
// (StiffGIPC/GIPC.cu:8405-8535) for the set-up at the start of a step, and
// _calFrictionGradient(_gd) / _calFrictionHessian(_gd) (GIPC.cu:1400-1845,
// 5090-5313) with the static-friction clamps of StiffGIPC/FrictionUtils.cuh
// (C1 for the pairs, C0 for the ground).  One thread per pair.  Every pair
// type takes one form here: relDX = sum_a w_a (x_a - x^t_a) with the pair's
// vertex weights, u = B^T relDX in the 3 x 2 tangent basis B, gradient_a =
// mu lambda w_a B f1(|u|) u / |u| and Hessian_ab = mu lambda w_a w_b B PD(M2)
// B^T, written to [PT+EE x 16 | PE x 9 | PP x 4] by MatIndex (the reference
// numbers the blocks with atomic counters) with the lower blocks moved to
// (col, row) and zeroed.  The ground's tangent projector is I - n n^T, which
// for the reference's y-up ground gives its xz entries.
#include "friction.hpp"
#include "wave_ops.hpp"

namespace sgipc {

namespace {

struct F3 {
    double x, y, z;
};
__device__ __forceinline__ F3 ld(const double* p, long i) { return F3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b)
{
    return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ F3 unit(F3 a)
{
    const double n = sqrt(dot(a, a));
    return F3{a.x / n, a.y / n, a.z / n};
}

// vertex ids, count and type (2 PP, 3 PE, 4 PT, 5 EE); -1: smooth encoding
__device__ __forceinline__ int decode(int4 pr, int* vid, int& type)
{
    if (pr.x >= 0) {
        if (pr.w < 0) return -1;
        vid[0] = pr.x; vid[1] = pr.y; vid[2] = pr.z; vid[3] = pr.w;
        type = 5;
        return 4;
    }
    vid[0] = -pr.x - 1;
    if (pr.y < 0) return -1;
    vid[1] = pr.y;
    if (pr.z < 0) {
        type = 2;
        return 2;
    }
    vid[2] = pr.z;
    if (pr.w < 0) {
        type = 3;
        return 3;
    }
    vid[3] = pr.w;
    type = 4;
    return 4;
}

__device__ __forceinline__ double sqdist(int type, const F3* P)
{
    if (type == 2) {
        const F3 d = sub(P[0], P[1]);
        return dot(d, d);
    }
    if (type == 3) {
        const F3 c = cross(sub(P[1], P[0]), sub(P[2], P[0])), e = sub(P[2], P[1]);
        return dot(c, c) / dot(e, e);
    }
    if (type == 4) {
        const F3 nv = cross(sub(P[2], P[1]), sub(P[3], P[1]));
        const double s = dot(sub(P[0], P[1]), nv);
        return s * s / dot(nv, nv);
    }
    const F3 nv = cross(sub(P[1], P[0]), sub(P[3], P[2]));
    const double s = dot(sub(P[2], P[0]), nv);
    return s * s / dot(nv, nv);
}

__global__ void k_friction_prepare(int n, const double* __restrict__ xt, const int4* __restrict__ pairs, double dh,
                                   double kappa, double* __restrict__ lambda, double* __restrict__ coord,
                                   double* __restrict__ basis, int* __restrict__ bad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int vid[4], type;
    const int k = decode(pairs[p], vid, type);
    if (k < 0) {
        atomicOr(bad, 1);
        return;
    }
    F3 P[4];
    for (int a = 0; a < k; ++a) P[a] = ld(xt, vid[a]);
    F3 b0, b1;
    double c0 = 0.0, c1 = 0.0;
    if (type == 4) {  // FrictionUtils.cuh:53-86
        const F3 e1 = sub(P[2], P[1]), e2 = sub(P[3], P[1]), w = sub(P[0], P[1]);
        const double A = dot(e1, e1), Bm = dot(e1, e2), C = dot(e2, e2), r0 = dot(e1, w), r1 = dot(e2, w);
        const double det = A * C - Bm * Bm;
        c0 = (C * r0 - Bm * r1) / det;
        c1 = (A * r1 - Bm * r0) / det;
        b0 = unit(e1);
        b1 = unit(cross(cross(e1, e2), e1));
    } else if (type == 5) {  // :157-196
        const F3 e20 = sub(P[0], P[2]), e01 = sub(P[1], P[0]), e23 = sub(P[3], P[2]);
        const double m00 = dot(e01, e01), m01 = -dot(e23, e01), m11 = dot(e23, e23);
        const double r0 = -dot(e20, e01), r1 = dot(e20, e23), det = m00 * m11 - m01 * m01;
        c0 = (m11 * r0 - m01 * r1) / det;
        c1 = (m00 * r1 - m01 * r0) / det;
        b0 = unit(e01);
        b1 = unit(cross(cross(e01, e23), e01));
    } else if (type == 3) {  // :251-272
        const F3 e12 = sub(P[2], P[1]), w = sub(P[0], P[1]);
        c0 = dot(w, e12) / dot(e12, e12);
        b0 = unit(e12);
        b1 = unit(cross(e12, w));
    } else {  // :301-331
        const F3 v01 = sub(P[1], P[0]);
        const F3 xc{0.0, -v01.z, v01.y}, yc{v01.z, 0.0, -v01.x};
        const F3 ch = dot(xc, xc) > dot(yc, yc) ? xc : yc;
        b0 = unit(ch);
        b1 = unit(cross(v01, ch));
    }
    coord[2L * p] = c0;
    coord[2L * p + 1] = c1;
    double* B = basis + 6L * p;
    B[0] = b0.x; B[1] = b0.y; B[2] = b0.z;
    B[3] = b1.x; B[4] = b1.y; B[5] = b1.z;
    // GIPC.cu:8519-8526 (RANK 2)
    const double d = sqdist(type, P), L = log(d / dh);
    lambda[p] = -kappa * 2.0 * sqrt(d) * (L * L * (2.0 * d - 2.0 * dh) + 2.0 * L * (d - dh) * (d - dh) / d);
}

// symmetric 2x2 eigen clamp
__device__ __forceinline__ void clamp_pd2(double M[2][2])
{
    const double a = M[0][0], b = 0.5 * (M[0][1] + M[1][0]), d = M[1][1];
    const double tr = 0.5 * (a + d), df = 0.5 * (a - d), rad = sqrt(df * df + b * b);
    double l0 = tr + rad, l1 = tr - rad;
    double v0x = 1.0, v0y = 0.0;
    if (rad > 0.0) {
        if (df >= 0.0) {
            v0x = df + rad;
            v0y = b;
        } else {
            v0x = b;
            v0y = rad - df;
        }
        const double nn = sqrt(v0x * v0x + v0y * v0y);
        v0x /= nn;
        v0y /= nn;
    }
    const double v1x = -v0y, v1y = v0x;
    l0 = fmax(l0, 0.0);
    l1 = fmax(l1, 0.0);
    M[0][0] = l0 * v0x * v0x + l1 * v1x * v1x;
    M[0][1] = l0 * v0x * v0y + l1 * v1x * v1y;
    M[1][0] = M[0][1];
    M[1][1] = l0 * v0y * v0y + l1 * v1y * v1y;
}

__global__ __launch_bounds__(128) void k_friction(int n, const double* __restrict__ x, const double* __restrict__ xt,
                                                  const int4* __restrict__ pairs, const int* __restrict__ mat_index,
                                                  int n4, int n3, const double* __restrict__ lambda,
                                                  const double* __restrict__ coord, const double* __restrict__ basis,
                                                  double eps2, double mu, int hess_offset, double* __restrict__ grad,
                                                  double* __restrict__ tvals, int* __restrict__ trows,
                                                  int* __restrict__ tcols, int* __restrict__ bad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int vid[4], type;
    const int k = decode(pairs[p], vid, type);
    if (k < 0) {
        atomicOr(bad, 1);
        return;
    }
    const double c0 = coord[2L * p], c1 = coord[2L * p + 1];
    double w[4] = {0.0, 0.0, 0.0, 0.0};
    if (type == 4) {
        w[0] = 1.0; w[1] = -1.0 + c0 + c1; w[2] = -c0; w[3] = -c1;
    } else if (type == 5) {
        w[0] = 1.0 - c0; w[1] = c0; w[2] = c1 - 1.0; w[3] = -c1;
    } else if (type == 3) {
        w[0] = 1.0; w[1] = c0 - 1.0; w[2] = -c0;
    } else {
        w[0] = 1.0; w[1] = -1.0;
    }
    F3 r{0.0, 0.0, 0.0};
    for (int a = 0; a < k; ++a) {
        const F3 dx = sub(ld(x, vid[a]), ld(xt, vid[a]));
        r.x += w[a] * dx.x;
        r.y += w[a] * dx.y;
        r.z += w[a] * dx.z;
    }
    const double* Bp = basis + 6L * p;
    const F3 B0{Bp[0], Bp[1], Bp[2]}, B1{Bp[3], Bp[4], Bp[5]};
    const double u0 = dot(B0, r), u1 = dot(B1, r);
    const double uu = u0 * u0 + u1 * u1, un = sqrt(uu), eps = sqrt(eps2);
    const double s = mu * lambda[p];
    if (grad != nullptr) {
        const double gs = s * (uu > eps2 ? 1.0 / un : (-un + 2.0 * eps) / (eps * eps));
        const F3 g3{gs * (B0.x * u0 + B1.x * u1), gs * (B0.y * u0 + B1.y * u1), gs * (B0.z * u0 + B1.z * u1)};
        for (int a = 0; a < k; ++a) {
            atomic_add(grad + 3L * vid[a], w[a] * g3.x);
            atomic_add(grad + 3L * vid[a] + 1, w[a] * g3.y);
            atomic_add(grad + 3L * vid[a] + 2, w[a] * g3.z);
        }
    }
    const double u[2] = {u0, u1};
    double M[2][2];
    if (uu > eps2) {
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) M[i][j] = (i == j ? 1.0 / un : 0.0) - u[i] * u[j] / (uu * un);
    } else {
        const double f1d = (-un + 2.0 * eps) / (eps * eps), f2 = 2.0 * (eps - un) / (eps * eps);
        const bool aniso = f2 != f1d && uu != 0.0;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) M[i][j] = (i == j ? f1d : 0.0) - (aniso ? (f1d - f2) * u[i] * u[j] / uu : 0.0);
    }
    clamp_pd2(M);
    const double Bc[2][3] = {{B0.x, B0.y, B0.z}, {B1.x, B1.y, B1.z}};
    double H3[3][3];
    for (int rr = 0; rr < 3; ++rr)
        for (int cc = 0; cc < 3; ++cc) {
            double acc = 0.0;
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j) acc += Bc[i][rr] * M[i][j] * Bc[j][cc];
            H3[rr][cc] = s * acc;
        }
    const int mi = mat_index[p];
    const long base = k == 4 ? 16L * mi : (k == 3 ? 16L * n4 + 9L * mi : 16L * n4 + 9L * n3 + 4L * mi);
    for (int ii = 0; ii < k; ++ii)
        for (int jj = 0; jj < k; ++jj) {
            const long kk = base + ii * k + jj;
            const int row = vid[ii] + hess_offset, col = vid[jj] + hess_offset;
            double* dst = tvals + 9 * kk;
            if (row <= col) {
                trows[kk] = row;
                tcols[kk] = col;
                const double ww = w[ii] * w[jj];
                for (int rr = 0; rr < 3; ++rr)
                    for (int cc = 0; cc < 3; ++cc) dst[cc * 3 + rr] = ww * H3[rr][cc];
            } else {
                trows[kk] = col;
                tcols[kk] = row;
                for (int e = 0; e < 9; ++e) dst[e] = 0.0;
            }
        }
}

__global__ void k_ground_prepare(int n, const double* __restrict__ xt, const int* __restrict__ verts,
                                 const double* __restrict__ nrm, double offset, double dh, double kappa,
                                 double* __restrict__ lambda)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const F3 nv{nrm[0], nrm[1], nrm[2]};
    const double dist = dot(nv, ld(xt, verts[i])) - offset;
    const double d2 = dist * dist, t = d2 - dh;
    const double gb = t * log(d2 / dh) * -2.0 - (t * t) / d2;
    lambda[i] = -kappa * 2.0 * sqrt(d2) * gb;
}

__global__ void k_ground_friction(int n, const double* __restrict__ x, const double* __restrict__ xt,
                                  const int* __restrict__ verts, const double* __restrict__ nrm,
                                  const double* __restrict__ lambda, double eps2, double mu, int hess_offset,
                                  double* __restrict__ grad, double* __restrict__ tvals, int* __restrict__ trows,
                                  int* __restrict__ tcols)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = verts[i];
    const F3 nv{nrm[0], nrm[1], nrm[2]};
    const F3 dx = sub(ld(x, v), ld(xt, v));
    const double dn = dot(dx, nv);
    const double V[3] = {dx.x - nv.x * dn, dx.y - nv.y * dn, dx.z - nv.z * dn};
    const double vv = V[0] * V[0] + V[1] * V[1] + V[2] * V[2], vn = sqrt(vv), eps = sqrt(eps2);
    const double m = mu * lambda[i];
    if (grad != nullptr) {
        const double gs = m / (vv > eps2 ? vn : eps);
        for (int d = 0; d < 3; ++d) atomic_add(grad + 3L * v + d, V[d] * gs);
    }
    const double n3[3] = {nv.x, nv.y, nv.z};
    double* t = tvals + 9L * i;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            const double P = (r == c ? 1.0 : 0.0) - n3[r] * n3[c];
            t[c * 3 + r] = vv > eps2 ? m * (P / vn - V[r] * V[c] / (vv * vn)) : m / eps * P;
        }
    trows[i] = v + hess_offset;
    tcols[i] = v + hess_offset;
}

inline int g128(long n) { return (int)((n + 127) / 128); }

}  // namespace

void launch_friction_prepare(hipStream_t s, int n, const double* xt, const int* pairs, double dhat_sq, double kappa,
                             double* lambda, double* coord, double* basis, int* bad)
{
    if (n <= 0) return;
    k_friction_prepare<<<g128(n), 128, 0, s>>>(n, xt, reinterpret_cast<const int4*>(pairs), dhat_sq, kappa, lambda,
                                               coord, basis, bad);
    SGIPC_LAUNCH_CHECK();
}

void launch_friction(hipStream_t s, int n, const double* x, const double* xt, const int* pairs, const int* mat_index,
                     int n4, int n3, const double* lambda, const double* coord, const double* basis, double eps2,
                     double mu, int hess_offset, double* grad, double* tvals, int* trows, int* tcols, int* bad)
{
    if (n <= 0) return;
    k_friction<<<g128(n), 128, 0, s>>>(n, x, xt, reinterpret_cast<const int4*>(pairs), mat_index, n4, n3, lambda,
                                       coord, basis, eps2, mu, hess_offset, grad, tvals, trows, tcols, bad);
    SGIPC_LAUNCH_CHECK();
}

void launch_ground_friction_prepare(hipStream_t s, int n, const double* xt, const int* verts, const double* normal,
                                    double offset, double dhat_sq, double kappa, double* lambda)
{
    if (n <= 0) return;
    k_ground_prepare<<<g128(n), 128, 0, s>>>(n, xt, verts, normal, offset, dhat_sq, kappa, lambda);
    SGIPC_LAUNCH_CHECK();
}

void launch_ground_friction(hipStream_t s, int n, const double* x, const double* xt, const int* verts,
                            const double* normal, const double* lambda, double eps2, double mu, int hess_offset,
                            double* grad, double* tvals, int* trows, int* tcols)
{
    if (n <= 0) return;
    k_ground_friction<<<g128(n), 128, 0, s>>>(n, x, xt, verts, normal, lambda, eps2, mu, hess_offset, grad, tvals,
                                              trows, tcols);
    SGIPC_LAUNCH_CHECK();
}

}  // namespace sgipc
