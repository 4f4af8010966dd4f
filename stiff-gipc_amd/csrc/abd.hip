// Affine-body (ABD) part of the hybrid system for gfx950.
//
//  * body term (StiffGIPC/abd_system/abd_system_function/setup_abd_system_
//    gradient_and_hessian.cu:200-289): H = M + PSD(kappa vol dt^2 d2V/da2)
//    on the affine block, G = M (q - q~) + kappa vol dt^2 dV/da, 10 upper
//    blocks per body at rows 4b..4b+3; fixed bodies keep H = M, G = 0.  The
//    orthogonality energy V(a) (abd_system/details/abd_energy.inl) is
//    sum (a_i.a_i - 1)^2 + 2 sum_{i<j} (a_i.a_j)^2; its 9x9 Hessian is
//    projected with a cyclic Jacobi eigen-solve (one body per thread: a
//    scene has hundreds of bodies, not millions).
//  * contact classes (GIPC.cu:10080-10256 partition, setup_abd_system_
//    gradient_and_hessian.cu:87-141,424-558 transforms): each point-level
//    upper contact block (i, j, H) becomes 1 FEM-FEM block, 4 ABD-FEM blocks
//    J_i^T H, or 16 ABD-ABD blocks J_i^T H J_j (blocks below the diagonal
//    zeroed at (col,row); a same-body pair i != j stands for (i,j) and (j,i),
//    so it contributes X + X^T); FEM-ABD blocks carry nothing.  Output is
//    class by class, each in input order (three scans).
//  * block-Jacobi (setup_abd_system_gradient_and_hessian.cu:560-655): the
//    inverse of every body's 12x12 diagonal block, Gauss-Jordan with partial
//    pivoting; apply z_b = P_b r_b with r.z partials.
#include <cstring>

#include <rocprim/device/device_scan.hpp>

#include "abd.hpp"
#include "wave_ops.hpp"

namespace sgipc {

namespace {

inline int g256(long n) { return (int)((n + 255) / 256); }

__device__ void mass12(const double* __restrict__ m10, double M[12][12])
{
#pragma unroll
    for (int i = 0; i < 12; ++i)
#pragma unroll
        for (int j = 0; j < 12; ++j) M[i][j] = 0.0;
    const double D[3][3] = {{m10[4], m10[5], m10[6]}, {m10[5], m10[7], m10[8]}, {m10[6], m10[8], m10[9]}};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        M[r][r] = m10[0];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            M[r][3 + 3 * r + c] = m10[1 + c];
            M[3 + 3 * r + c][r] = m10[1 + c];
#pragma unroll
            for (int d = 0; d < 3; ++d) M[3 + 3 * r + c][3 + 3 * r + d] = D[c][d];
        }
    }
}

// in place: A (n x n) -> V diag(max(l, 0)) V^T
template <int N>
__device__ void psd_project(double A[N][N])
{
    double V[N][N];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) V[i][j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
                const double a = A[i][j] * A[i][j];
                tot += a;
                if (i != j) off += a;
            }
        if (off <= 1e-32 * tot || off == 0.0) break;
        for (int p = 0; p < N - 1; ++p)
            for (int q = p + 1; q < N; ++q) {
                const double apq = A[p][q];
                if (apq == 0.0) continue;
                const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < N; ++k) {
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < N; ++k) {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < N; ++k) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    double l[N];
    for (int k = 0; k < N; ++k) l[k] = A[k][k] > 0.0 ? A[k][k] : 0.0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            double s = 0.0;
            for (int k = 0; k < N; ++k) s += V[i][k] * l[k] * V[j][k];
            A[i][j] = s;
        }
}

__device__ __forceinline__ double dot3(const double* a, const double* b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Motor body (setup_abd_system_gradient_and_hessian.cu:271-339): the target
// q_p keeps q_tilde's rows a2, a3 and adds A~ (R e_y - e_y) and A~ (R e_z - e_z)
// to their second and third entries (R the rotation by theta = motor_speed dt
// about the body x axis; the reference reaches this through J^-1 J_delta over
// the rest points 0, e_x, e_y, e_z), each row normalised; then, with
// W = strength * M[6:12, 6:12] (the body mass), G += W (q - q_p) on the a2, a3
// rows and H += W.
__device__ void abd_motor(const double* __restrict__ qb, const double* __restrict__ qt, double theta,
                          const double W[6][6], double H[12][12], double G[12])
{
    const double c = cos(theta), s = sin(theta);
    const double d2[3] = {0.0, c - 1.0, s};   // R e_y - e_y
    const double d3[3] = {0.0, -s, c - 1.0};  // R e_z - e_z
    double dq[6];
    for (int r = 1; r < 3; ++r) {
        const double* ar = qt + 3 + 3 * r;  // row r of A~
        double ap[3] = {ar[0], ar[1] + (ar[0] * d2[0] + ar[1] * d2[1] + ar[2] * d2[2]),
                        ar[2] + (ar[0] * d3[0] + ar[1] * d3[1] + ar[2] * d3[2])};
        const double nrm = sqrt(ap[0] * ap[0] + ap[1] * ap[1] + ap[2] * ap[2]);
        for (int k = 0; k < 3; ++k) dq[3 * (r - 1) + k] = qb[3 + 3 * r + k] - ap[k] / nrm;
    }
    for (int i = 0; i < 6; ++i) {
        double acc = 0.0;
        for (int j = 0; j < 6; ++j) acc += W[i][j] * dq[j];
        G[6 + i] += acc;
        for (int j = 0; j < 6; ++j) H[6 + i][6 + j] += W[i][j];
    }
}

__global__ __launch_bounds__(64) void k_abd_body(int nb, const double* __restrict__ q, const double* __restrict__ qt,
                                                 const double* __restrict__ mass10, const double* __restrict__ vol,
                                                 double kappa, double dt, const int* __restrict__ fixed,
                                                 double motor_speed, double motor_strength,
                                                 double* __restrict__ grad, double* __restrict__ tvals,
                                                 int* __restrict__ trows, int* __restrict__ tcols)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const double dt2 = dt * dt;
    double H[12][12];
    mass12(mass10 + 10 * (long)b, H);
    double G[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) G[i] = 0.0;
    const int btype = fixed ? fixed[b] : 0;  // 0 free, 1 fixed, 2 motor (BodyBoundaryType)
    double W[6][6];
    if (btype == 2)
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) W[i][j] = motor_strength * H[6 + i][6 + j];
    if (btype != 1) {
        double qb[12], dq[12];
        for (int i = 0; i < 12; ++i) {
            qb[i] = q[12 * (long)b + i];
            dq[i] = qb[i] - qt[12 * (long)b + i];
        }
        for (int i = 0; i < 12; ++i) {
            double s = 0.0;
            for (int j = 0; j < 12; ++j) s += H[i][j] * dq[j];
            G[i] = s;
        }
        const double kvt2 = kappa * vol[b] * dt2;
        const double* a[3] = {qb + 3, qb + 6, qb + 9};
        double S[9][9];
        for (int i = 0; i < 3; ++i) {
            for (int d = 0; d < 3; ++d) {
                double v = 4.0 * (dot3(a[i], a[i]) - 1.0) * a[i][d];
                for (int j = 0; j < 3; ++j)
                    if (j != i) v += 4.0 * dot3(a[j], a[i]) * a[j][d];
                G[3 + 3 * i + d] += kvt2 * v;
            }
            for (int j = 0; j < 3; ++j)
                for (int c = 0; c < 3; ++c)
                    for (int d = 0; d < 3; ++d) {
                        double v;
                        if (i == j) {
                            v = 8.0 * a[i][c] * a[i][d] + (c == d ? 4.0 * (dot3(a[i], a[i]) - 1.0) : 0.0);
                            for (int k = 0; k < 3; ++k)
                                if (k != i) v += 4.0 * a[k][c] * a[k][d];
                        } else {
                            v = 4.0 * a[j][c] * a[i][d] + (c == d ? 4.0 * dot3(a[i], a[j]) : 0.0);
                        }
                        S[3 * i + c][3 * j + d] = kvt2 * v;
                    }
        }
        psd_project<9>(S);
        for (int i = 0; i < 9; ++i)
            for (int j = 0; j < 9; ++j) H[3 + i][3 + j] += S[i][j];
        if (btype == 2) abd_motor(qb, qt + 12 * (long)b, motor_speed * dt, W, H, G);
    }
    if (grad)
        for (int i = 0; i < 12; ++i) grad[12 * (long)b + i] = G[i];
    int kk = 0;
    for (int ii = 0; ii < 4; ++ii)
        for (int jj = ii; jj < 4; ++jj, ++kk) {
            const long o = 10 * (long)b + kk;
            trows[o] = 4 * b + ii;
            tcols[o] = 4 * b + jj;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) tvals[9 * o + c * 3 + r] = H[3 * ii + r][3 * jj + c];
        }
}

// class of a point-level contact block: 0 FEM-FEM, 1 ABD-FEM, 2 ABD-ABD, 3 none
__device__ __forceinline__ int contact_class(int i, int j, int n_abd)
{
    const bool ia = i < n_abd, ja = j < n_abd;
    if (!ia && !ja) return 0;
    if (ia && !ja) return 1;
    if (ia && ja) return 2;
    return 3;
}

__global__ void k_contact_count(long n, const int* __restrict__ rows, const int* __restrict__ cols, int n_abd,
                                int* __restrict__ c0, int* __restrict__ c1, int* __restrict__ c2)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > n) return;
    int cls = k < n ? contact_class(rows[k], cols[k], n_abd) : 3;
    c0[k] = cls == 0 ? 1 : 0;
    c1[k] = cls == 1 ? 4 : 0;
    c2[k] = cls == 2 ? 16 : 0;
}

__global__ __launch_bounds__(128) void k_contact_map(long n, const double* __restrict__ vals,
                                                     const int* __restrict__ rows, const int* __restrict__ cols,
                                                     int n_abd, const int* __restrict__ point_body,
                                                     const double* __restrict__ x_bar, int nb,
                                                     const int* __restrict__ body_fixed,
                                                     const int* __restrict__ s0, const int* __restrict__ s1,
                                                     const int* __restrict__ s2, const int* __restrict__ fem_btype,
                                                     double* __restrict__ ov, int* __restrict__ orow,
                                                     int* __restrict__ ocol)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = rows[k], j = cols[k];
    const int cls = contact_class(i, j, n_abd);
    if (cls == 3) return;
    const int fem0 = 4 * nb - n_abd;
    double Hm[3][3];  // row-major
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Hm[r][c] = vals[9 * k + c * 3 + r];
    const int base1 = s0[n], base2 = s0[n] + s1[n];
    if (cls == 0) {
        // FEM-FEM: zeroed at Dirichlet points (GIPC.cu:10332-10363)
        const long o = s0[k];
        orow[o] = i + fem0;
        ocol[o] = j + fem0;
        const bool zero = fem_btype && (fem_btype[i - n_abd] != 0 || fem_btype[j - n_abd] != 0);
#pragma unroll
        for (int e = 0; e < 9; ++e) ov[9 * o + e] = zero ? 0.0 : vals[9 * k + e];
        return;
    }
    if (cls == 1) {
        const int bi = point_body[i];
        // ABD-FEM: zeroed for a fixed body or a Dirichlet FEM point
        // (setup_abd_system_gradient_and_hessian.cu:564-568)
        const bool zero = (body_fixed && body_fixed[bi] == 1) || (fem_btype && fem_btype[j - n_abd] != 0);
        const double xb[3] = {x_bar[3 * (long)i], x_bar[3 * (long)i + 1], x_bar[3 * (long)i + 2]};
        const long o0 = base1 + s1[k];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const long o = o0 + a;
            orow[o] = 4 * bi + a;
            ocol[o] = j + fem0;
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double v = a == 0 ? Hm[r][c] : xb[r] * Hm[a - 1][c];
                    ov[9 * o + c * 3 + r] = zero ? 0.0 : v;
                }
        }
        return;
    }
    // ABD-ABD: X = J_i^T H J_j; (a, b) block of X over [p | a1 | a2 | a3]
    const int bi = point_body[i], bj = point_body[j];
    const bool zero = body_fixed && (body_fixed[bi] == 1 || body_fixed[bj] == 1);
    const bool sym = bi == bj && i != j;
    const double x[4] = {1.0, x_bar[3 * (long)i], x_bar[3 * (long)i + 1], x_bar[3 * (long)i + 2]};
    const double y[4] = {1.0, x_bar[3 * (long)j], x_bar[3 * (long)j + 1], x_bar[3 * (long)j + 2]};
    // X[(3a + r), (3b + c)]: a = 0 -> coordinate rows of J (r is the row index of H);
    // a = 1 + s -> x_r * H(s, .).  Generally X_ab(r, c) = u_a(r) H v_b(c) with
    // u_0(r) = e_r, u_{1+s}(r) = x_r e_s and v likewise.
    const long o0 = base2 + s2[k];
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
            const long o = o0 + 4 * a + b;
            const int row = 4 * bi + a, col = 4 * bj + b;
            const bool keep = row <= col && !zero;
            orow[o] = row <= col ? row : col;
            ocol[o] = row <= col ? col : row;
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    double v = 0.0;
                    if (keep) {
                        // X_ab(r, c) = sum_{p,q} J_i(p, 3a+r) H(p,q) J_j(q, 3b+c)
                        // J(p, 3a+r) = (a == 0) ? [p == r] : [p == a-1] * x_r
                        const int pi = a == 0 ? r : a - 1;
                        const double wi = a == 0 ? 1.0 : x[1 + r];
                        const int qj = b == 0 ? c : b - 1;
                        const double wj = b == 0 ? 1.0 : y[1 + c];
                        v = wi * Hm[pi][qj] * wj;
                        if (sym) {
                            // + X^T_ab(r, c) = X_ba(c, r)
                            const int pi2 = b == 0 ? c : b - 1;
                            const double wi2 = b == 0 ? 1.0 : x[1 + c];
                            const int qj2 = a == 0 ? r : a - 1;
                            const double wj2 = a == 0 ? 1.0 : y[1 + r];
                            v += wi2 * Hm[pi2][qj2] * wj2;
                        }
                    }
                    ov[9 * o + c * 3 + r] = v;
                }
        }
}

// Inverse of each body's 12x12 diagonal block (Gauss-Jordan, partial pivoting).
__global__ __launch_bounds__(64) void k_abd_jacobi_build(const double* __restrict__ vals,
                                                         const int* __restrict__ rows,
                                                         const int* __restrict__ cols, int nnz, int nb,
                                                         double* __restrict__ inv, int* __restrict__ bad)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double A[12][12], I[12][12];
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) {
            A[i][j] = 0.0;
            I[i][j] = i == j ? 1.0 : 0.0;
        }
    // blocks of rows 4b..4b+3 (rows sorted): lower_bound of 4b
    int lo = 0, hi = nnz;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (rows[mid] < 4 * b) lo = mid + 1; else hi = mid;
    }
    for (int k = lo; k < nnz && rows[k] < 4 * b + 4; ++k) {
        const int r = rows[k], c = cols[k];
        if (c >= 4 * b + 4) continue;
        const int ra = r - 4 * b, ca = c - 4 * b;
        for (int u = 0; u < 3; ++u)
            for (int v = 0; v < 3; ++v) {
                const double h = vals[9 * (long)k + v * 3 + u];
                A[3 * ra + u][3 * ca + v] = h;
                if (r != c) A[3 * ca + v][3 * ra + u] = h;
            }
    }
    bool ok = true;
    for (int c = 0; c < 12; ++c) {
        int pr = c;
        for (int r = c + 1; r < 12; ++r)
            if (fabs(A[r][c]) > fabs(A[pr][c])) pr = r;
        if (A[pr][c] == 0.0) {
            ok = false;
            break;
        }
        if (pr != c)
            for (int k = 0; k < 12; ++k) {
                double t = A[c][k];
                A[c][k] = A[pr][k];
                A[pr][k] = t;
                t = I[c][k];
                I[c][k] = I[pr][k];
                I[pr][k] = t;
            }
        const double piv = 1.0 / A[c][c];
        for (int k = 0; k < 12; ++k) {
            A[c][k] *= piv;
            I[c][k] *= piv;
        }
        for (int r = 0; r < 12; ++r) {
            if (r == c) continue;
            const double f = A[r][c];
            if (f == 0.0) continue;
            for (int k = 0; k < 12; ++k) {
                A[r][k] -= f * A[c][k];
                I[r][k] -= f * I[c][k];
            }
        }
    }
    if (!ok) atomicOr(bad, 1);
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) inv[144 * (long)b + 12 * i + j] = ok ? I[i][j] : 0.0;
}

// z_b = P_b r_b for the body rows; r.z partial per workgroup.  One thread per
// (body, output row).  Returns at once when ctl->done is set.
template <typename TV>
__global__ __launch_bounds__(256) void k_abd_jacobi_apply(const double* __restrict__ inv, int nb,
                                                          const TV* __restrict__ r, TV* __restrict__ z,
                                                          double* __restrict__ rz_part, const PcgCtl* __restrict__ ctl)
{
    __shared__ double red[16];
    if (ctl != nullptr && ctl->done) return;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    double rzv = 0.0;
    if (t < 12L * nb) {
        const long b = t / 12;
        const int i = (int)(t % 12);
        for (int j = 0; j < 12; ++j) acc += inv[144 * b + 12 * i + j] * (double)r[12 * b + j];
        z[t] = (TV)acc;
        rzv = (double)r[t] * (double)(TV)acc;
    }
    if (rz_part) {
        const double s = block_sum(rzv, red);
        if (threadIdx.x == 0) rz_part[blockIdx.x] = s;
    }
}

// Energy gradient of the contact points into the system rows
// (setup_abd_system_gradient_and_hessian.cu:349-385, fem_linear_subsystem.cu:
// 56-84): a body point adds J^T g to its body's 12 rows (not for a fixed
// body), J^T g = [g, xbar_0 g, xbar_1 g, xbar_2 g] over [p | a1 | a2 | a3]
// (a_r row r of A: rows 3 + 3 r + c take g_r xbar_c); a FEM point adds g to
// its vertex rows, which follow the 4 nb body rows.  Atomic like the
// reference: the order of the additions is not fixed.
__global__ void k_point_gradient(int n_points, int n_abd, const double* __restrict__ gp,
                                 const int* __restrict__ point_body, const double* __restrict__ x_bar,
                                 const int* __restrict__ body_fixed, int nb, double* __restrict__ grad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_points) return;
    const double g[3] = {gp[3L * i], gp[3L * i + 1], gp[3L * i + 2]};
    if (i >= n_abd) {
        double* d = grad + 3L * (4L * nb + (i - n_abd));
        for (int r = 0; r < 3; ++r) atomicAdd(d + r, g[r]);
        return;
    }
    const int b = point_body[i];
    if (body_fixed && body_fixed[b] == 1) return;
    double* d = grad + 12L * b;
    for (int r = 0; r < 3; ++r) {
        atomicAdd(d + r, g[r]);
        for (int c = 0; c < 3; ++c) atomicAdd(d + 3 + 3 * r + c, g[r] * x_bar[3L * i + c]);
    }
}

// Dirichlet vertices (GIPC.cu:10444-10462): a triplet with an end on a FEM
// row (>= row0) whose vertex is fixed is zeroed; body rows are left alone.
__global__ void k_dirichlet_zero(long n, const int* __restrict__ rows, const int* __restrict__ cols, int row0,
                                 const int* __restrict__ btype, double* __restrict__ vals)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int r = rows[k], c = cols[k];
    const bool zero = (r >= row0 && btype[r - row0] != 0) || (c >= row0 && btype[c - row0] != 0);
    if (zero)
        for (int e = 0; e < 9; ++e) vals[9 * k + e] = 0.0;
}

// gradient rows of Dirichlet vertices are zero (fem_linear_subsystem.cu:73-76)
__global__ void k_dirichlet_zero_rows(int n, const int* __restrict__ btype, double* __restrict__ g)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n && btype[v] != 0)
        for (int r = 0; r < 3; ++r) g[3L * v + r] = 0.0;
}

}  // namespace

void launch_abd_body(hipStream_t s, int nb, const double* q, const double* q_tilde, const double* mass10,
                     const double* vol, double kappa, double dt, const int* fixed, double motor_speed,
                     double motor_strength, double* grad, double* tvals, int* trows, int* tcols)
{
    if (nb <= 0) return;
    k_abd_body<<<(nb + 63) / 64, 64, 0, s>>>(nb, q, q_tilde, mass10, vol, kappa, dt, fixed, motor_speed,
                                              motor_strength, grad, tvals, trows, tcols);
    SGIPC_LAUNCH_CHECK();
}

void launch_point_gradient(hipStream_t s, int n_points, int n_abd, const double* gp, const int* point_body,
                           const double* x_bar, const int* body_fixed, int nb, double* grad)
{
    if (n_points <= 0) return;
    k_point_gradient<<<g256(n_points), 256, 0, s>>>(n_points, n_abd, gp, point_body, x_bar, body_fixed, nb, grad);
    SGIPC_LAUNCH_CHECK();
}

void launch_dirichlet_zero(hipStream_t s, long n, const int* rows, const int* cols, int row0, const int* btype,
                           double* vals)
{
    if (n <= 0 || btype == nullptr) return;
    k_dirichlet_zero<<<g256(n), 256, 0, s>>>(n, rows, cols, row0, btype, vals);
    SGIPC_LAUNCH_CHECK();
}

void launch_dirichlet_zero_rows(hipStream_t s, int n, const int* btype, double* g)
{
    if (n <= 0 || btype == nullptr) return;
    k_dirichlet_zero_rows<<<g256(n), 256, 0, s>>>(n, btype, g);
    SGIPC_LAUNCH_CHECK();
}

long ContactMapper::count(hipStream_t s, const int* rows, const int* cols, long n, int n_abd)
{
    n_ = n;
    if (n <= 0) return 0;
    for (int c = 0; c < 3; ++c) {
        cnt_[c].reserve(n + 1);
        scan_[c].reserve(n + 1);
    }
    k_contact_count<<<g256(n + 1), 256, 0, s>>>(n, rows, cols, n_abd, cnt_[0].get(), cnt_[1].get(), cnt_[2].get());
    SGIPC_LAUNCH_CHECK();
    size_t need = 0;
    SGIPC_HIP(rocprim::exclusive_scan(nullptr, need, cnt_[0].get(), scan_[0].get(), 0, (size_t)n + 1,
                                      rocprim::plus<int>(), s));
    temp_.reserve(need);
    for (int c = 0; c < 3; ++c)
        SGIPC_HIP(rocprim::exclusive_scan(temp_.get(), need, cnt_[c].get(), scan_[c].get(), 0, (size_t)n + 1,
                                          rocprim::plus<int>(), s));
    int h[3] = {0, 0, 0};
    for (int c = 0; c < 3; ++c)
        SGIPC_HIP(hipMemcpyAsync(&h[c], scan_[c].get() + n, sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    total_ = (long)h[0] + h[1] + h[2];
    return total_;
}

void ContactMapper::write(hipStream_t s, const double* vals, const int* rows, const int* cols, int n_abd,
                          const int* point_body, const double* x_bar, int nb, const int* body_fixed,
                          const int* fem_btype, double* ov, int* orow, int* ocol, long capacity)
{
    if (n_ <= 0 || total_ == 0) return;
    SGIPC_CHECK(total_ <= capacity, SGIPC_CAPACITY, "contact map exceeds the output capacity");
    k_contact_map<<<(int)((n_ + 127) / 128), 128, 0, s>>>(n_, vals, rows, cols, n_abd, point_body, x_bar, nb,
                                                          body_fixed, scan_[0].get(), scan_[1].get(), scan_[2].get(),
                                                          fem_btype, ov, orow, ocol);
    SGIPC_LAUNCH_CHECK();
}

void launch_abd_jacobi_build(hipStream_t s, const double* vals, const int* rows, const int* cols, int nnz, int nb,
                             double* inv, int* bad)
{
    if (nb <= 0) return;
    k_abd_jacobi_build<<<(nb + 63) / 64, 64, 0, s>>>(vals, rows, cols, nnz, nb, inv, bad);
    SGIPC_LAUNCH_CHECK();
}

template <typename TV>
int launch_abd_jacobi_apply(hipStream_t s, const double* inv, int nb, const TV* r, TV* z, double* rz_part,
                            const PcgCtl* ctl)
{
    if (nb <= 0) return 0;
    const int g = g256(12L * nb);
    k_abd_jacobi_apply<TV><<<g, 256, 0, s>>>(inv, nb, r, z, rz_part, ctl);
    SGIPC_LAUNCH_CHECK();
    return rz_part ? g : 0;
}
template int launch_abd_jacobi_apply<double>(hipStream_t, const double*, int, const double*, double*, double*,
                                             const PcgCtl*);
template int launch_abd_jacobi_apply<float>(hipStream_t, const double*, int, const float*, float*, double*,
                                            const PcgCtl*);

void AbdJacobi::build(hipStream_t s, const double* vals, const int* rows, const int* cols, int nnz, int nb)
{
    n_bodies = nb;
    if (nb <= 0) return;
    inv.reserve(144L * nb);
    bad_.reserve(1);
    SGIPC_HIP(hipMemsetAsync(bad_.get(), 0, sizeof(int), s));
    launch_abd_jacobi_build(s, vals, rows, cols, nnz, nb, inv.get(), bad_.get());
    int h = 0;
    SGIPC_HIP(hipMemcpyAsync(&h, bad_.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    SGIPC_CHECK(h == 0, SGIPC_NOT_SPD, "singular affine-body block");
}

template <typename TV>
int AbdJacobi::apply(hipStream_t s, const TV* r, TV* z, double* rz_part, const PcgCtl* ctl) const
{
    return launch_abd_jacobi_apply<TV>(s, inv.get(), n_bodies, r, z, rz_part, ctl);
}
template int AbdJacobi::apply<double>(hipStream_t, const double*, double*, double*, const PcgCtl*) const;
template int AbdJacobi::apply<float>(hipStream_t, const float*, float*, double*, const PcgCtl*) const;

int AbdJacobi::partials() const { return n_bodies > 0 ? g256(12L * n_bodies) : 0; }

}  // namespace sgipc
