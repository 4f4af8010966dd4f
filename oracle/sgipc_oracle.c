/*
 * CPU ORACLE -- TEST INFRASTRUCTURE ONLY (see sgipc_oracle.h).
 *
 * Plain C restatement of the reference's Newton linear-solve path.  Each
 * function cites the reference file:line it follows.  The oracle is written for
 * clarity first; OpenMP parallel loops keep the full-size scenes tractable when
 * it is timed as the CPU baseline.
 */
#include "sgipc_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_fault = 0;
void orc_set_fault(int fault) { g_fault = fault; }
int orc_get_fault(void) { return g_fault; }

/* ------------------------------------------------------------------------- */
/* small dense helpers                                                        */
/* ------------------------------------------------------------------------- */
static double det3(const double F[3][3])
{
    return F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) -
           F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
           F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
}

static int levi(int i, int j, int k)
{
    if (i == j || j == k || i == k) return 0;
    /* even permutations of (0,1,2) */
    if ((i == 0 && j == 1 && k == 2) || (i == 1 && j == 2 && k == 0) || (i == 2 && j == 0 && k == 1))
        return 1;
    return -1;
}

/* Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (row-major).
 * On return A holds the eigenvalues on its diagonal, V (row-major) the
 * eigenvectors as columns. */
static void jacobi_eig(int n, double* A, double* V)
{
    for (int i = 0; i < n * n; ++i) V[i] = 0.0;
    for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double a = A[i * n + j] * A[i * n + j];
                tot += a;
                if (i != j) off += a;
            }
        if (off <= 1e-32 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = A[p * n + q];
                if (apq == 0.0) continue;
                double app = A[p * n + p], aqq = A[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
}

/* ------------------------------------------------------------------------- */
/* Stable Neo-Hookean tet                                                     */
/* StiffGIPC/femEnergy.cu:432-459 (PK1), 1641-1717 (projected Hessian),       */
/* 1719-1857 (kernel), 1389-1420 (dF/dx), 324-350 (Ds).                       */
/* Psi(F) = mu/2 (|F|^2 - 3) - mu (J - 1) + lam/2 (J - 1)^2 with mu = lengthRate,  */
/* lam = volumeRate; P = mu F + (lam (J-1) - mu) dJ/dF.                        */
/* vec(F) is column-major: q = 3*j + i  <->  F(i,j).                          */
/* ------------------------------------------------------------------------- */
static void tet_F(const double* xs /* 12 */, const double* dminv /* 9 rm */, double F[3][3])
{
    double Ds[3][3];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) Ds[i][k] = xs[3 * (k + 1) + i] - xs[i];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += Ds[i][k] * dminv[3 * k + j];
            F[i][j] = s;
        }
}

double orc_tet_snh_energy(const double* x12, const double* dm_inv, double mu, double lam)
{
    double F[3][3];
    tet_F(x12, dm_inv, F);
    double I2 = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) I2 += F[i][j] * F[i][j];
    double J = det3(F);
    return 0.5 * mu * (I2 - 3.0) - mu * (J - 1.0) + 0.5 * lam * (J - 1.0) * (J - 1.0);
}

static void snh_hessian9(const double F[3][3], double mu, double lam, double H[81])
{
    double J = det3(F);
    double g[9];
    /* cofactor = dJ/dF */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l)
                    for (int m = 0; m < 3; ++m)
                        for (int n = 0; n < 3; ++n)
                            s += 0.5 * levi(i, k, m) * levi(j, l, n) * F[k][l] * F[m][n];
            g[3 * j + i] = s;
        }
    double c = lam * (J - 1.0) - mu;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) {
                    int q = 3 * j + i, r = 3 * l + k;
                    double hj = 0;
                    for (int m = 0; m < 3; ++m)
                        for (int n = 0; n < 3; ++n) hj += levi(i, k, m) * levi(j, l, n) * F[m][n];
                    H[q * 9 + r] = (q == r ? mu : 0.0) + lam * g[q] * g[r] + c * hj;
                }
}

static void project_psd9(const double H[81], double P[81])
{
    double A[81], V[81];
    memcpy(A, H, sizeof(A));
    /* symmetrise against round-off before the eigen-solve */
    for (int i = 0; i < 9; ++i)
        for (int j = i + 1; j < 9; ++j) {
            double s = 0.5 * (A[i * 9 + j] + A[j * 9 + i]);
            A[i * 9 + j] = A[j * 9 + i] = s;
        }
    jacobi_eig(9, A, V);
    double lam[9];
    for (int i = 0; i < 9; ++i) lam[i] = A[i * 9 + i] > 0.0 ? A[i * 9 + i] : 0.0;
    for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) {
            double s = 0;
            for (int k = 0; k < 9; ++k) s += V[i * 9 + k] * lam[k] * V[j * 9 + k];
            P[i * 9 + j] = s;
        }
}

void orc_snh_dpdf(const double* F9, double mu, double lam, double* H81, double* Hproj81)
{
    double F[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) F[i][j] = F9[3 * j + i];
    double H[81];
    snh_hessian9(F, mu, lam, H);
    if (H81) memcpy(H81, H, sizeof(H));
    if (Hproj81) project_psd9(H, Hproj81);
}

/* upper-block writer of femEnergy.cu:1832-1856: block (ii,jj), ii<=jj local;
 * if the global row > col the transposed block is stored at (col,row). */
static void write_upper_blocks(int nv, const int* gid, const double* H /* (3nv)^2 rm */,
                               double* tvals, int* trows, int* tcols)
{
    int n3 = 3 * nv, kk = 0;
    for (int ii = 0; ii < nv; ++ii)
        for (int jj = ii; jj < nv; ++jj, ++kk) {
            int row = gid[ii], col = gid[jj];
            double* v = tvals + 9 * (size_t)kk;
            if (row <= col) {
                trows[kk] = row; tcols[kk] = col;
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) v[c * 3 + r] = H[(3 * ii + r) * n3 + 3 * jj + c];
            } else {
                trows[kk] = col; tcols[kk] = row;
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) v[c * 3 + r] = H[(3 * jj + r) * n3 + 3 * ii + c];
            }
        }
}

int orc_tet_snh(int n_tets, const double* x, const int* tets, const double* dm_inv,
                const double* vol, const double* mu, const double* lam, double dt,
                int hess_offset, double* grad, double* tvals, int* trows, int* tcols)
{
#pragma omp parallel for schedule(static)
    for (int t = 0; t < n_tets; ++t) {
        const int* tv = tets + 4 * (size_t)t;
        double xs[12];
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 3; ++d) xs[3 * a + d] = x[3 * (size_t)tv[a] + d];
        const double* Dmi = dm_inv + 9 * (size_t)t;
        double F[3][3];
        tet_F(xs, Dmi, F);
        double Hq[81], Hp[81];
        snh_hessian9(F, mu[t], lam[t], Hq);
        project_psd9(Hq, Hp);
        /* D(a,j) = dF(:,j)/dx_a coefficient */
        double D[4][3];
        for (int j = 0; j < 3; ++j) {
            D[0][j] = -(Dmi[0 * 3 + j] + Dmi[1 * 3 + j] + Dmi[2 * 3 + j]);
            for (int a = 1; a < 4; ++a) D[a][j] = Dmi[(a - 1) * 3 + j];
        }
        double scale = dt * dt * vol[t];
        if (grad) {
            double J = det3(F);
            double P[3][3], cof[3][3];
            cof[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
            cof[0][1] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
            cof[0][2] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
            cof[1][0] = F[2][1] * F[0][2] - F[2][2] * F[0][1];
            cof[1][1] = F[2][2] * F[0][0] - F[2][0] * F[0][2];
            cof[1][2] = F[2][0] * F[0][1] - F[2][1] * F[0][0];
            cof[2][0] = F[0][1] * F[1][2] - F[1][1] * F[0][2];
            cof[2][1] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
            cof[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
            double c = lam[t] * (J - 1.0) - mu[t];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) P[i][j] = mu[t] * F[i][j] + c * cof[i][j];
            for (int a = 0; a < 4; ++a)
                for (int d = 0; d < 3; ++d) {
                    double s = 0;
                    for (int j = 0; j < 3; ++j) s += D[a][j] * P[d][j];
                    s *= scale;
#pragma omp atomic
                    grad[3 * (size_t)tv[a] + d] += s;
                }
        }
        double H12[144];
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 3; ++d)
                for (int b = 0; b < 4; ++b)
                    for (int e = 0; e < 3; ++e) {
                        double s = 0;
                        for (int j = 0; j < 3; ++j)
                            for (int jp = 0; jp < 3; ++jp)
                                s += D[a][j] * D[b][jp] * Hp[(3 * j + d) * 9 + 3 * jp + e];
                        H12[(3 * a + d) * 12 + 3 * b + e] = scale * s;
                    }
        int gid[4] = {tv[0] + hess_offset, tv[1] + hess_offset, tv[2] + hess_offset,
                      tv[3] + hess_offset};
        write_upper_blocks(4, gid, H12, tvals + 90 * (size_t)t, trows + 10 * (size_t)t,
                           tcols + 10 * (size_t)t);
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Baraff-Witkin cloth triangle (stretch + shear)                             */
/* StiffGIPC/femEnergy.cu:2146-2215 (kernel), 351-395 (dE/dF), 535-629 and    */
/* 631-700 (projected stretch / shear Hessians), 967-1005 (energy),           */
/* 1171-1210 (dF/dx), 304-322 (Ds).  F = Ds Dm^-1 is 3x2 with columns         */
/* f_u, f_v; vec(F) is column-major (q = 3*c + r).                            */
/*   E = area * ( k_st * sum_{w=u,v} psi(|f_w|) + k_sh * I6^2 ),              */
/*   psi(s) = (s-1)^2 + [s>1] rate (s-1)^3,  I6 = f_u . f_v.                  */
/* The Hessian is projected term by term: each stretch term onto             */
/* psi'' f^ f^T + max(psi'/s, 0)(I - f^ f^T), the shear term onto the         */
/* non-negative part of its eigen-decomposition.                             */
/* ------------------------------------------------------------------------- */
static void tri_F(const double* xs /* 9 */, const double* dminv /* 4 rm */, double F[3][2])
{
    for (int i = 0; i < 3; ++i) {
        double d1 = xs[3 + i] - xs[i], d2 = xs[6 + i] - xs[i];
        for (int j = 0; j < 2; ++j) F[i][j] = d1 * dminv[j] + d2 * dminv[2 + j];
    }
}

static void stretch_terms(double sig, double rate, double* dpsi, double* d2psi)
{
    double e = sig - 1.0;
    *dpsi = 2.0 * e + (sig > 1.0 ? 3.0 * rate * e * e : 0.0);
    *d2psi = 2.0 + (sig > 1.0 ? 6.0 * rate * e : 0.0);
}

double orc_tri_bw_energy(const double* x9, const double* dm_inv2, double area, double k_st,
                         double k_sh, double rate)
{
    double F[3][2];
    tri_F(x9, dm_inv2, F);
    double E = 0.0;
    for (int c = 0; c < 2; ++c) {
        double sig = sqrt(F[0][c] * F[0][c] + F[1][c] * F[1][c] + F[2][c] * F[2][c]);
        double e = sig - 1.0;
        E += k_st * (e * e + (sig * sig > 1.0 ? rate * e * e * e : 0.0));
    }
    double I6 = F[0][0] * F[0][1] + F[1][0] * F[1][1] + F[2][0] * F[2][1];
    return area * (E + k_sh * I6 * I6);
}

/* dE/dF (per unit area) and the projected 6x6 d2E/dF2 (row-major over vec(F)). */
static void bw_pepf_hq(const double F[3][2], double k_st, double k_sh, double rate, double P[3][2],
                       double Hq[36])
{
    memset(Hq, 0, 36 * sizeof(double));
    for (int c = 0; c < 2; ++c) {
        double f[3] = {F[0][c], F[1][c], F[2][c]};
        double sig = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        double d1, d2;
        stretch_terms(sig, rate, &d1, &d2);
        double fh[3] = {f[0] / sig, f[1] / sig, f[2] / sig};
        for (int r = 0; r < 3; ++r) P[r][c] = k_st * d1 * fh[r];
        double iso = d1 / sig > 0.0 ? d1 / sig : 0.0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                Hq[(3 * c + a) * 6 + 3 * c + b] +=
                    k_st * ((a == b ? iso : 0.0) + (d2 - iso) * fh[a] * fh[b]);
    }
    double I6 = F[0][0] * F[0][1] + F[1][0] * F[1][1] + F[2][0] * F[2][1];
    for (int r = 0; r < 3; ++r) {
        P[r][0] += k_sh * 2.0 * I6 * F[r][1];
        P[r][1] += k_sh * 2.0 * I6 * F[r][0];
    }
    /* shear: d2(I6^2)/dF2 = 2 g g^T + 2 I6 S, g = (f_v, f_u), S swaps the columns */
    double g[6] = {F[0][1], F[1][1], F[2][1], F[0][0], F[1][0], F[2][0]};
    double A[36], V[36];
    for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) A[a * 6 + b] = 2.0 * g[a] * g[b] + ((a % 3 == b % 3 && a / 3 != b / 3) ? 2.0 * I6 : 0.0);
    jacobi_eig(6, A, V);
    for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) {
            double sum = 0.0;
            for (int k = 0; k < 6; ++k) {
                double l = A[k * 6 + k];
                if (l > 0.0 || g_fault == 4) sum += V[a * 6 + k] * l * V[b * 6 + k];
            }
            Hq[a * 6 + b] += k_sh * sum;
        }
}

void orc_bw_dpdf(const double* F6, double k_st, double k_sh, double rate, double* P6, double* Hproj36)
{
    double F[3][2], P[3][2], H[36];
    for (int c = 0; c < 2; ++c)
        for (int r = 0; r < 3; ++r) F[r][c] = F6[3 * c + r];
    bw_pepf_hq(F, k_st, k_sh, rate, P, H);
    if (P6)
        for (int c = 0; c < 2; ++c)
            for (int r = 0; r < 3; ++r) P6[3 * c + r] = P[r][c];
    if (Hproj36) memcpy(Hproj36, H, sizeof(H));
}

int orc_tri_bw(int n_tris, const double* x, const int* tris, const double* dm_inv2, const double* area,
               double k_st, double k_sh, double rate, double dt, int hess_offset, double* grad,
               double* tvals, int* trows, int* tcols)
{
#pragma omp parallel for schedule(static)
    for (int t = 0; t < n_tris; ++t) {
        const int* tv = tris + 3 * (size_t)t;
        double xs[9];
        for (int a = 0; a < 3; ++a)
            for (int d = 0; d < 3; ++d) xs[3 * a + d] = x[3 * (size_t)tv[a] + d];
        const double* Dmi = dm_inv2 + 4 * (size_t)t;
        double F[3][2], P[3][2], Hq[36];
        tri_F(xs, Dmi, F);
        bw_pepf_hq(F, k_st, k_sh, rate, P, Hq);
        /* D[a][j]: dF(:,j)/dx_a */
        double D[3][2];
        for (int j = 0; j < 2; ++j) {
            D[0][j] = -(Dmi[j] + Dmi[2 + j]);
            D[1][j] = Dmi[j];
            D[2][j] = Dmi[2 + j];
        }
        double scale = dt * dt * area[t];
        if (grad)
            for (int a = 0; a < 3; ++a)
                for (int d = 0; d < 3; ++d) {
                    double g = scale * (D[a][0] * P[d][0] + D[a][1] * P[d][1]);
#pragma omp atomic
                    grad[3 * (size_t)tv[a] + d] += g;
                }
        double H9[81];
        for (int a = 0; a < 3; ++a)
            for (int d = 0; d < 3; ++d)
                for (int b = 0; b < 3; ++b)
                    for (int e = 0; e < 3; ++e) {
                        double sum = 0.0;
                        for (int j = 0; j < 2; ++j)
                            for (int jp = 0; jp < 2; ++jp)
                                sum += D[a][j] * D[b][jp] * Hq[(3 * j + d) * 6 + 3 * jp + e];
                        H9[(3 * a + d) * 9 + 3 * b + e] = scale * sum;
                    }
        int gid[3] = {tv[0] + hess_offset, tv[1] + hess_offset, tv[2] + hess_offset};
        write_upper_blocks(3, gid, H9, tvals + 54 * (size_t)t, trows + 6 * (size_t)t,
                           tcols + 6 * (size_t)t);
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Quadratic (constant-Q) bending, StiffGIPC/femEnergy.cu:2490-2618 (Q),      */
/* 2620-2754 (kernel).  Vertex order [e0, e1, opposite of tri 1, opposite of  */
/* tri 2]; Q = 3/(A0+A1) K K^T with cotangent weights K; E = coef/2 x^T Q x   */
/* per coordinate, coef = |rest e0 - rest e1| k dt^2.  Boundary edges          */
/* (adj[1] == -1) and degenerate hinges get Q = 0; the kernel then writes     */
/* ten zero blocks at (0, 0).                                                 */
/* ------------------------------------------------------------------------- */
static double cot3(const double* a, const double* b)
{
    double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    double cn = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    if (cn < 1e-12) return 0.0;
    return (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) / cn;
}

void orc_quad_bending_q(int n_edges, const double* rest_x, const int* edges, const int* adj, double* Q16)
{
    for (int i = 0; i < n_edges; ++i) {
        double* Q = Q16 + 16 * (size_t)i;
        memset(Q, 0, 16 * sizeof(double));
        if (adj[2 * i + 1] == -1) continue;
        int vid[4] = {edges[2 * i], edges[2 * i + 1], adj[2 * i], adj[2 * i + 1]};
        double v[4][3];
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 3; ++d) v[a][d] = rest_x[3 * (size_t)vid[a] + d];
        double e0[3], e1[3], e2[3], e3[3], e4[3], m0[3];
        for (int d = 0; d < 3; ++d) {
            e0[d] = v[1][d] - v[0][d];
            e1[d] = v[2][d] - v[0][d];
            e2[d] = v[3][d] - v[0][d];
            e3[d] = v[2][d] - v[1][d];
            e4[d] = v[3][d] - v[1][d];
            m0[d] = -e0[d];
        }
        double c01[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0]};
        double c02[3] = {e0[1] * e2[2] - e0[2] * e2[1], e0[2] * e2[0] - e0[0] * e2[2], e0[0] * e2[1] - e0[1] * e2[0]};
        double n01 = sqrt(c01[0] * c01[0] + c01[1] * c01[1] + c01[2] * c01[2]);
        double n02 = sqrt(c02[0] * c02[0] + c02[1] * c02[1] + c02[2] * c02[2]);
        if (n01 < 1e-12 || n02 < 1e-12) continue;
        double A = 0.5 * n01 + 0.5 * n02;
        if (A < 1e-12) continue;
        double k01 = cot3(e0, e1), k02 = cot3(e0, e2), k03 = cot3(m0, e3), k04 = cot3(m0, e4);
        double K[4] = {k03 + k04, k01 + k02, -k01 - k03, -k02 - k04};
        int bad = 0;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) {
                Q[r * 4 + c] = 3.0 / A * K[r] * K[c];
                if (!isfinite(Q[r * 4 + c])) bad = 1;
            }
        if (bad) memset(Q, 0, 16 * sizeof(double));
    }
}

int orc_quad_bending(int n_edges, const double* x, const double* rest_x, const int* edges, const int* adj,
                     const double* Q16, double k_bend, double dt, int hess_offset, double* grad,
                     double* tvals, int* trows, int* tcols)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n_edges; ++i) {
        double* tv = tvals + 90 * (size_t)i;
        int* tr = trows + 10 * (size_t)i;
        int* tc = tcols + 10 * (size_t)i;
        if (adj[2 * i + 1] == -1) {
            memset(tv, 0, 90 * sizeof(double));
            for (int k = 0; k < 10; ++k) tr[k] = tc[k] = 0;
            continue;
        }
        int vid[4] = {edges[2 * i], edges[2 * i + 1], adj[2 * i], adj[2 * i + 1]};
        const double* Q = Q16 + 16 * (size_t)i;
        double L = 0.0;
        for (int d = 0; d < 3; ++d) {
            double e = rest_x[3 * (size_t)vid[0] + d] - rest_x[3 * (size_t)vid[1] + d];
            L += e * e;
        }
        double coef = sqrt(L) * k_bend * dt * dt;
        if (grad)
            for (int a = 0; a < 4; ++a)
                for (int d = 0; d < 3; ++d) {
                    double g = 0.0;
                    for (int b = 0; b < 4; ++b) g += Q[a * 4 + b] * x[3 * (size_t)vid[b] + d];
                    g *= coef;
#pragma omp atomic
                    grad[3 * (size_t)vid[a] + d] += g;
                }
        double H12[144];
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 3; ++d)
                for (int b = 0; b < 4; ++b)
                    for (int e = 0; e < 3; ++e) H12[(3 * a + d) * 12 + 3 * b + e] = d == e ? coef * Q[a * 4 + b] : 0.0;
        int gid[4] = {vid[0] + hess_offset, vid[1] + hess_offset, vid[2] + hess_offset, vid[3] + hess_offset};
        write_upper_blocks(4, gid, H12, tv, tr, tc);
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Affine body dynamics (ABD)                                                 */
/* StiffGIPC/abd_system/abd_system_function/setup_abd_system_gradient_and_    */
/* hessian.cu:200-289 (body term), 87-141 (J^T H J), 424-558 (contact classes) */
/* , 560-655 (block-Jacobi); abd_system/details/abd_energy.inl (shape energy),*/
/* details/abd_jacobi_matrix.inl (J).  q = [p, a1, a2, a3]; a point with rest */
/* coordinates xb sits at x = p + (xb.a1, xb.a2, xb.a3), J = [I | e_r xb^T].   */
/* mass10 = (m, m xb (3), m xb xb^T upper (xx, xy, xz, yy, yz, zz)) summed     */
/* over the body's mass points; M = sum m J^T J.                               */
/* ------------------------------------------------------------------------- */
static void abd_mass12(const double* m10, double M[144])
{
    memset(M, 0, 144 * sizeof(double));
    double mm = m10[0];
    const double* mx = m10 + 1;
    double D[3][3] = {{m10[4], m10[5], m10[6]}, {m10[5], m10[7], m10[8]}, {m10[6], m10[8], m10[9]}};
    for (int r = 0; r < 3; ++r) M[r * 12 + r] = mm;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            M[r * 12 + 3 + 3 * r + c] = mx[c];
            M[(3 + 3 * r + c) * 12 + r] = mx[c];
        }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            for (int d = 0; d < 3; ++d) M[(3 + 3 * r + c) * 12 + 3 + 3 * r + d] = D[c][d];
}

double orc_abd_shape_energy(const double* q)
{
    const double *a1 = q + 3, *a2 = q + 6, *a3 = q + 9;
#define D3(u, v) ((u)[0] * (v)[0] + (u)[1] * (v)[1] + (u)[2] * (v)[2])
    double e = 0.0, t;
    t = D3(a1, a1) - 1.0; e += t * t;
    t = D3(a2, a2) - 1.0; e += t * t;
    t = D3(a3, a3) - 1.0; e += t * t;
    t = D3(a1, a2); e += 2.0 * t * t;
    t = D3(a2, a3); e += 2.0 * t * t;
    t = D3(a3, a1); e += 2.0 * t * t;
    return e;
}

/* gradient (9) and unprojected Hessian (9x9 row-major) over (a1, a2, a3) */
void orc_abd_shape_grad_hess(const double* q, double* g9, double* H81)
{
    const double* a[3] = {q + 3, q + 6, q + 9};
    for (int i = 0; i < 3; ++i) {
        const double* ai = a[i];
        for (int d = 0; d < 3; ++d) {
            double v = 4.0 * (D3(ai, ai) - 1.0) * ai[d];
            for (int j = 0; j < 3; ++j)
                if (j != i) v += 4.0 * D3(a[j], ai) * a[j][d];
            if (g9) g9[3 * i + d] = v;
        }
    }
    if (!H81) return;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int c = 0; c < 3; ++c)
                for (int d = 0; d < 3; ++d) {
                    double v;
                    if (i == j) {
                        v = 8.0 * a[i][c] * a[i][d] + (c == d ? 4.0 * (D3(a[i], a[i]) - 1.0) : 0.0);
                        for (int k = 0; k < 3; ++k)
                            if (k != i) v += 4.0 * a[k][c] * a[k][d];
                    } else {
                        /* d2/dai daj of 2 (ai.aj)^2 = 4 (aj ai^T + (ai.aj) I) */
                        v = 4.0 * a[j][c] * a[i][d] + (c == d ? 4.0 * D3(a[i], a[j]) : 0.0);
                    }
                    H81[(3 * i + c) * 9 + 3 * j + d] = v;
                }
}

static void project_psd_n(int n, double* H)
{
    double A[144], V[144];
    memcpy(A, H, sizeof(double) * n * n);
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double sm = 0.5 * (A[i * n + j] + A[j * n + i]);
            A[i * n + j] = A[j * n + i] = sm;
        }
    jacobi_eig(n, A, V);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double sm = 0.0;
            for (int k = 0; k < n; ++k)
                if (A[k * n + k] > 0.0) sm += V[i * n + k] * A[k * n + k] * V[j * n + k];
            H[i * n + j] = sm;
        }
}

static int inv12(double* A);

/* ABDJacobi{xb}.to_mat() (abd_jacobi_matrix.inl:32-43): row p has 1 at column
 * p and xb^T at columns 3 + 3p .. 5 + 3p; written into rows r0..r0+2 of a
 * 12x12 row-major matrix */
static void abd_jacobi_rows(const double* xb, double* J, int r0)
{
    for (int p = 0; p < 3; ++p) {
        double* row = J + 12 * (r0 + p);
        for (int c = 0; c < 12; ++c) row[c] = 0.0;
        row[p] = 1.0;
        for (int c = 0; c < 3; ++c) row[3 + 3 * p + c] = xb[c];
    }
}

/* Motor body term (setup_abd_system_gradient_and_hessian.cu:271-339), step
 * by step as there: J over the rest points 0, e_x, e_y, e_z; J_delta over
 * 0, 0, R e_y - e_y, R e_z - e_z (R: angle theta about x); q_p = J^-1 J_delta
 * q~ + q~ with segments 6 and 9 normalised; dq = q - q_p with its first six
 * entries zeroed; W = strength * M[6:12, 6:12]; G += W dq, H += W. */
static void abd_motor_term(const double* qb, const double* qt, const double* M, double theta, double strength,
                           double* H, double* G)
{
    const double bar[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const double c = cos(theta), s = sin(theta);
    const double R[3][3] = {{1, 0, 0}, {0, c, -s}, {0, s, c}};
    double J[144], Jd[144];
    for (int k = 0; k < 4; ++k) abd_jacobi_rows(bar[k], J, 3 * k);
    const double zero[3] = {0, 0, 0};
    abd_jacobi_rows(zero, Jd, 0);
    abd_jacobi_rows(zero, Jd, 3);
    for (int k = 2; k < 4; ++k) {
        double d[3];
        for (int i = 0; i < 3; ++i) d[i] = R[i][0] * bar[k][0] + R[i][1] * bar[k][1] + R[i][2] * bar[k][2] - bar[k][i];
        abd_jacobi_rows(d, Jd, 3 * k);
    }
    inv12(J);
    double t[12], qp[12];
    for (int i = 0; i < 12; ++i) {
        double sm = 0.0;
        for (int j = 0; j < 12; ++j) sm += Jd[i * 12 + j] * qt[j];
        t[i] = sm;
    }
    for (int i = 0; i < 12; ++i) {
        double sm = 0.0;
        for (int j = 0; j < 12; ++j) sm += J[i * 12 + j] * t[j];
        qp[i] = sm + qt[i];
    }
    for (int seg = 6; seg <= 9; seg += 3) {
        double nrm = sqrt(qp[seg] * qp[seg] + qp[seg + 1] * qp[seg + 1] + qp[seg + 2] * qp[seg + 2]);
        for (int k = 0; k < 3; ++k) qp[seg + k] /= nrm;
    }
    double dq[12];
    for (int i = 0; i < 12; ++i) dq[i] = i < 6 ? 0.0 : qb[i] - qp[i];
    for (int i = 6; i < 12; ++i) {
        double sm = 0.0;
        for (int j = 6; j < 12; ++j) {
            const double w = strength * M[i * 12 + j];
            sm += w * dq[j];
            H[i * 12 + j] += w;
        }
        G[i] += sm;
    }
}

/* fixed[b] is the reference's BodyBoundaryType (body_boundary_type.h:3-8):
 * 0 free, 1 fixed, 2 motor. */
int orc_abd_body(int n_bodies, const double* q, const double* q_tilde, const double* mass10, const double* volume,
                 double kappa, double dt, const int* fixed, double motor_speed, double motor_strength, double* grad,
                 double* tvals, int* trows, int* tcols)
{
#pragma omp parallel for schedule(static)
    for (int b = 0; b < n_bodies; ++b) {
        double M[144], H[144], G[12];
        abd_mass12(mass10 + 10 * (size_t)b, M);
        memcpy(H, M, sizeof(H));
        for (int i = 0; i < 12; ++i) G[i] = 0.0;
        const int bt = fixed ? fixed[b] : 0;
        if (bt != 1) {
            const double* qb = q + 12 * (size_t)b;
            const double* qt = q_tilde + 12 * (size_t)b;
            for (int i = 0; i < 12; ++i) {
                double sm = 0.0;
                for (int j = 0; j < 12; ++j) sm += M[i * 12 + j] * (qb[j] - qt[j]);
                G[i] = sm;
            }
            double kvt2 = kappa * volume[b] * dt * dt;
            double g9[9], H9[81];
            orc_abd_shape_grad_hess(qb, g9, H9);
            for (int i = 0; i < 81; ++i) H9[i] *= kvt2;
            project_psd_n(9, H9);
            for (int i = 0; i < 9; ++i) {
                G[3 + i] += kvt2 * g9[i];
                for (int j = 0; j < 9; ++j) H[(3 + i) * 12 + 3 + j] += H9[i * 9 + j];
            }
            if (bt == 2) abd_motor_term(qb, qt, M, motor_speed * dt, motor_strength, H, G);
        }
        if (grad)
            for (int i = 0; i < 12; ++i) grad[12 * (size_t)b + i] = G[i];
        int kk = 0;
        for (int ii = 0; ii < 4; ++ii)
            for (int jj = ii; jj < 4; ++jj, ++kk) {
                size_t o = 10 * (size_t)b + kk;
                trows[o] = 4 * b + ii;
                tcols[o] = 4 * b + jj;
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) tvals[9 * o + c * 3 + r] = H[(3 * ii + r) * 12 + 3 * jj + c];
            }
    }
    return 0;
}

/* J^T H J between points with rest coordinates x (left) and y (right); 12x12 row-major */
static void abd_jt_h_j(const double* x, const double* Hrm /* 3x3 row-major */, const double* y, double* out)
{
    double J1[36], J2[36];
    memset(J1, 0, sizeof(J1));
    memset(J2, 0, sizeof(J2));
    for (int r = 0; r < 3; ++r) {
        J1[r * 12 + r] = 1.0;
        J2[r * 12 + r] = 1.0;
        for (int c = 0; c < 3; ++c) {
            J1[r * 12 + 3 + 3 * r + c] = x[c];
            J2[r * 12 + 3 + 3 * r + c] = y[c];
        }
    }
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) {
            double sm = 0.0;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) sm += J1[a * 12 + i] * Hrm[a * 3 + b] * J2[b * 12 + j];
            out[i * 12 + j] = sm;
        }
}

/* fem_btype (NULL: none): per FEM point, nonzero = Dirichlet; FEM-FEM and
 * ABD-FEM blocks with such an end are zeroed (GIPC.cu:10332-10363;
 * setup_abd_system_gradient_and_hessian.cu:564-568). */
long orc_abd_contact_map(long n_trip, const double* vals, const int* rows, const int* cols, int n_abd_points,
                         const int* point_body, const double* x_bar, int n_bodies, const int* body_fixed,
                         const int* fem_btype, double* ovals, int* orows, int* ocols)
{
    long o = 0;
    const int fem0 = 4 * n_bodies - n_abd_points;  /* FEM point id -> system row */
    for (int cls = 0; cls < 3; ++cls)
        for (long k = 0; k < n_trip; ++k) {
            int i = rows[k], j = cols[k];
            int ia = i < n_abd_points, ja = j < n_abd_points;
            const double* v = vals + 9 * (size_t)k;
            double Hrm[9];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) Hrm[r * 3 + c] = v[c * 3 + r];
            if (cls == 0 && !ia && !ja) {
                orows[o] = i + fem0;
                ocols[o] = j + fem0;
                if (fem_btype && (fem_btype[i - n_abd_points] || fem_btype[j - n_abd_points]))
                    memset(ovals + 9 * (size_t)o, 0, 9 * sizeof(double));
                else
                    memcpy(ovals + 9 * (size_t)o, v, 9 * sizeof(double));
                ++o;
            } else if (cls == 1 && ia && !ja) {
                int bi = point_body[i];
                int zero = (body_fixed && body_fixed[bi] == 1) || (fem_btype && fem_btype[j - n_abd_points]);
                const double* xb = x_bar + 3 * (size_t)i;
                /* (J^T H) rows 3a..3a+2: a = 0 -> H, a = 1 + r -> xb (x) H row r */
                for (int a = 0; a < 4; ++a) {
                    double* d = ovals + 9 * (size_t)o;
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) {
                            double val = a == 0 ? Hrm[r * 3 + c] : xb[r] * Hrm[(a - 1) * 3 + c];
                            d[c * 3 + r] = zero ? 0.0 : val;
                        }
                    orows[o] = 4 * bi + a;
                    ocols[o] = j + fem0;
                    ++o;
                }
            } else if (cls == 2 && ia && ja) {
                int bi = point_body[i], bj = point_body[j];
                int zero = body_fixed && (body_fixed[bi] == 1 || body_fixed[bj] == 1);
                double X[144];
                abd_jt_h_j(x_bar + 3 * (size_t)i, Hrm, x_bar + 3 * (size_t)j, X);
                if (bi == bj && i != j) {
                    /* the stored upper vertex block stands for (i,j) and (j,i): X + X^T */
                    double T[144];
                    for (int a = 0; a < 12; ++a)
                        for (int b = 0; b < 12; ++b) T[a * 12 + b] = X[a * 12 + b] + X[b * 12 + a];
                    memcpy(X, T, sizeof(T));
                }
                for (int a = 0; a < 4; ++a)
                    for (int b = 0; b < 4; ++b) {
                        int row = 4 * bi + a, col = 4 * bj + b;
                        double* d = ovals + 9 * (size_t)o;
                        int keep = row <= col && !zero;
                        orows[o] = row <= col ? row : col;
                        ocols[o] = row <= col ? col : row;
                        for (int r = 0; r < 3; ++r)
                            for (int c = 0; c < 3; ++c) d[c * 3 + r] = keep ? X[(3 * a + r) * 12 + 3 * b + c] : 0.0;
                        ++o;
                    }
            }
        }
    return o;
}

/* 12x12 inverse by Gauss-Jordan with partial pivoting (row-major, in place) */
static int inv12(double* A)
{
    double I[144];
    int n = 12;
    memset(I, 0, sizeof(I));
    for (int i = 0; i < n; ++i) I[i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int pr = c;
        for (int r = c + 1; r < n; ++r)
            if (fabs(A[r * n + c]) > fabs(A[pr * n + c])) pr = r;
        if (A[pr * n + c] == 0.0) return -1;
        if (pr != c)
            for (int k = 0; k < n; ++k) {
                double t = A[c * n + k]; A[c * n + k] = A[pr * n + k]; A[pr * n + k] = t;
                t = I[c * n + k]; I[c * n + k] = I[pr * n + k]; I[pr * n + k] = t;
            }
        double piv = 1.0 / A[c * n + c];
        for (int k = 0; k < n; ++k) { A[c * n + k] *= piv; I[c * n + k] *= piv; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double f = A[r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) { A[r * n + k] -= f * A[c * n + k]; I[r * n + k] -= f * I[c * n + k]; }
        }
    }
    memcpy(A, I, sizeof(I));
    return 0;
}

int orc_abd_jacobi_build(long nnz, const double* vals, const int* rows, const int* cols, int n_bodies,
                         double* inv144)
{
    memset(inv144, 0, sizeof(double) * 144 * (size_t)n_bodies);
    for (long k = 0; k < nnz; ++k) {
        int r = rows[k], c = cols[k];
        if (r >= 4 * n_bodies || c >= 4 * n_bodies || r / 4 != c / 4) continue;
        double* P = inv144 + 144 * (size_t)(r / 4);
        const double* v = vals + 9 * (size_t)k;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                P[(3 * (r % 4) + a) * 12 + 3 * (c % 4) + b] = v[b * 3 + a];
                if (r != c) P[(3 * (c % 4) + b) * 12 + 3 * (r % 4) + a] = v[b * 3 + a];
            }
    }
    int bad = 0;
    for (int b = 0; b < n_bodies; ++b)
        if (inv12(inv144 + 144 * (size_t)b)) bad = 1;
    return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------- */
/* Point-triangle barrier, GIPC rank-2 variant                               */
/* StiffGIPC/GIPC.cu:4733-4984 (NEWF branch), mlbvh.cu:234-242 (_d_PT),       */
/* GIPC_PDerivative.cuh:2095-2291 (only column 8 of dF/dx is non-zero: the    */
/* gradient of s = dist/sqrt(dHat)).                                          */
/* E(I5) = kappa dHat^2 (I5-1)^2 ln(I5)^2 with I5 = s^2; lambda0 = d2E/ds2.   */
/* ------------------------------------------------------------------------- */
double orc_barrier_energy(double I5, double dhat_sq, double kappa)
{
    double L = log(I5);
    return kappa * dhat_sq * dhat_sq * (I5 - 1.0) * (I5 - 1.0) * L * L;
}

double orc_barrier_lambda0(double I5, double dhat_sq, double kappa)
{
    double L = log(I5);
    double lsign = (g_fault == 3) ? -1.0 : 1.0;
    return -(4.0 * kappa * dhat_sq * dhat_sq *
             (4.0 * I5 + lsign * L - 3.0 * I5 * I5 * L * L + 6.0 * I5 * L - 2.0 * I5 * I5 +
              I5 * L * L - 7.0 * I5 * I5 * L - 2.0)) /
           I5;
}

int orc_barrier_pt(int n_pairs, const double* x, const int* pairs, const double* pair_x,
                   double dhat_sq, double kappa, int hess_offset, double* grad,
                   double* tvals, int* trows, int* tcols)
{
    const double gass = 1e-6;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < n_pairs; ++p) {
        const int* pr = pairs + 4 * (size_t)p;
        int vid[4] = {-pr[0] - 1, pr[1], pr[2], pr[3]};
        double P[4][3];
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 3; ++d)
                P[a][d] = pair_x ? pair_x[12 * (size_t)p + 3 * a + d] : x[3 * (size_t)vid[a] + d];
        double e1[3], e2[3], w[3], b[3];
        for (int d = 0; d < 3; ++d) {
            e1[d] = P[2][d] - P[1][d];
            e2[d] = P[3][d] - P[1][d];
            w[d] = P[0][d] - P[1][d];
        }
        b[0] = e1[1] * e2[2] - e1[2] * e2[1];
        b[1] = e1[2] * e2[0] - e1[0] * e2[2];
        b[2] = e1[0] * e2[1] - e1[1] * e2[0];
        double bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
        double aTb = w[0] * b[0] + w[1] * b[1] + w[2] * b[2];
        double d2 = aTb * aTb / bb;
        double dis = sqrt(d2);
        double sq = sqrt(dhat_sq);
        double I5 = d2 / dhat_sq;
        double nb = sqrt(bb), sgn = aTb >= 0 ? 1.0 : -1.0;
        double nh[3] = {b[0] / nb, b[1] / nb, b[2] / nb};
        /* barycentric coordinates of the foot point */
        double c1[3], c2[3];
        c1[0] = w[1] * e2[2] - w[2] * e2[1];
        c1[1] = w[2] * e2[0] - w[0] * e2[2];
        c1[2] = w[0] * e2[1] - w[1] * e2[0];
        c2[0] = e1[1] * w[2] - e1[2] * w[1];
        c2[1] = e1[2] * w[0] - e1[0] * w[2];
        c2[2] = e1[0] * w[1] - e1[1] * w[0];
        double be1 = (c1[0] * b[0] + c1[1] * b[1] + c1[2] * b[2]) / bb;
        double be2 = (c2[0] * b[0] + c2[1] * b[1] + c2[2] * b[2]) / bb;
        double be0 = 1.0 - be1 - be2;
        double coef[4] = {1.0, -be0, -be1, -be2};
        double g[12];
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 3; ++d) g[3 * a + d] = sgn * coef[a] * nh[d] / sq;
        double Il = I5;
        if (d2 < gass * dhat_sq) Il = gass;
        double lam0 = orc_barrier_lambda0(Il, dhat_sq, kappa);
        if (grad) {
            double L = log(I5);
            double s = dis / sq;
            double pk = 2.0 * (2.0 * kappa * dhat_sq * dhat_sq * L * (I5 - 1.0) * (I5 + I5 * L - 1.0)) / I5;
            for (int a = 0; a < 4; ++a)
                for (int d = 0; d < 3; ++d) {
                    double v = g[3 * a + d] * s * pk;
#pragma omp atomic
                    grad[3 * (size_t)vid[a] + d] += v;
                }
        }
        double* tv = tvals + 144 * (size_t)p;
        int* tr = trows + 16 * (size_t)p;
        int* tc = tcols + 16 * (size_t)p;
        for (int ii = 0; ii < 4; ++ii)
            for (int jj = 0; jj < 4; ++jj) {
                int kk = ii * 4 + jj;
                int row = vid[ii] + hess_offset, col = vid[jj] + hess_offset;
                double* v = tv + 9 * kk;
                if (row <= col) {
                    tr[kk] = row; tc[kk] = col;
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) v[c * 3 + r] = lam0 * g[3 * ii + r] * g[3 * jj + c];
                } else {
                    tr[kk] = col; tc[kk] = row;
                    for (int e = 0; e < 9; ++e) v[e] = 0.0;
                }
            }
    }
    return 0;
}

void orc_dirichlet_zero(long n_trip, const int* rows, const int* cols, int row0, const int* btype, double* vals)
{
    for (long k = 0; k < n_trip; ++k) {
        const int r = rows[k], c = cols[k];
        if ((r >= row0 && btype[r - row0]) || (c >= row0 && btype[c - row0]))
            for (int e = 0; e < 9; ++e) vals[9 * k + e] = 0.0;
    }
}

int orc_mass_diag(int n, const double* mass, int hess_offset, double* tvals, int* trows,
                  int* tcols)
{
    for (int i = 0; i < n; ++i) {
        double* v = tvals + 9 * (size_t)i;
        for (int e = 0; e < 9; ++e) v[e] = 0.0;
        v[0] = v[4] = v[8] = mass[i];
        trows[i] = tcols[i] = i + hess_offset;
    }
    return 0;
}

/* OpenMP threads of the parallel loops (the CPU baseline's thread count).   */
void orc_set_threads(int k)
{
#ifdef _OPENMP
    if (k > 0) omp_set_num_threads(k);
#else
    (void)k;
#endif
}

/* ------------------------------------------------------------------------- */
/* Converter::srbk_convert -- StiffGIPC/linear_system/utils/converter.cu:32-51, */
/* 88-131, 171-238: radix sort on (row, col), reduce by key in sorted order   */
/* (blocks of more than 64 triplets: ordered sum of 64-triplet chunk sums).   */
/* ------------------------------------------------------------------------- */
/* Stable LSD radix sort of the triplet indices by key (8-bit digits over the
 * key's used bits; per-thread histograms, prefix in thread order so ties
 * keep index order): the order of a stable sort by (key, index). */
#define ORC_SEG_CHUNK 64

static void radix_sort_idx(long n, const uint64_t* keys, long* idx)
{
    uint64_t mx = 0;
    for (long i = 0; i < n; ++i)
        if (keys[i] > mx) mx = keys[i];
    int bits = 0;
    while (bits < 64 && (mx >> bits) != 0) ++bits;
    long* tmp = (long*)malloc(sizeof(long) * n);
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    long* hist = (long*)malloc(sizeof(long) * 256 * (size_t)nt);
    for (int sh = 0; sh < bits; sh += 8) {
#pragma omp parallel num_threads(nt)
        {
            int t = 0;
#ifdef _OPENMP
            t = omp_get_thread_num();
#endif
            long lo = n * t / nt, hi = n * (t + 1) / nt;
            long* h = hist + 256 * (size_t)t;
            for (int d = 0; d < 256; ++d) h[d] = 0;
            for (long i = lo; i < hi; ++i) ++h[(keys[idx[i]] >> sh) & 255];
#pragma omp barrier
#pragma omp single
            {
                long acc = 0;
                for (int d = 0; d < 256; ++d)
                    for (int q = 0; q < nt; ++q) {
                        long c = hist[256 * (size_t)q + d];
                        hist[256 * (size_t)q + d] = acc;
                        acc += c;
                    }
            }
            for (long i = lo; i < hi; ++i) tmp[h[(keys[idx[i]] >> sh) & 255]++] = idx[i];
        }
        memcpy(idx, tmp, sizeof(long) * n);
    }
    free(hist);
    free(tmp);
}

long orc_convert(long n_trip, const double* tvals, const int* trows, const int* tcols,
                 double* ovals, int* orows, int* ocols)
{
    if (n_trip <= 0) return 0;
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * n_trip);
    long* idx = (long*)malloc(sizeof(long) * n_trip);
    uint32_t mc = 0;
    for (long i = 0; i < n_trip; ++i)
        if ((uint32_t)tcols[i] > mc) mc = (uint32_t)tcols[i];
    int cb = 0;
    while (cb < 32 && (mc >> cb) != 0) ++cb;
    /* (row, col) order: row above the column's bits */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n_trip; ++i) {
        keys[i] = ((uint64_t)(uint32_t)trows[i] << cb) + (uint64_t)(uint32_t)tcols[i];
        idx[i] = i;
    }
    radix_sort_idx(n_trip, keys, idx);
    /* unique keys: heads of the sorted runs, then each run summed in order */
    long* head = (long*)malloc(sizeof(long) * (n_trip + 1));
    long u = 0;
    for (long s = 0; s < n_trip; ++s)
        if (s == 0 || keys[idx[s]] != keys[idx[s - 1]]) head[u++] = s;
    head[u] = n_trip;
    /* a block of more than ORC_SEG_CHUNK triplets is the ordered sum of its
     * ORC_SEG_CHUNK-triplet chunk sums (the library's fixed order) */
#pragma omp parallel for schedule(dynamic, 1024)
    for (long q = 0; q < u; ++q) {
        long i0 = idx[head[q]];
        orows[q] = trows[i0];
        ocols[q] = tcols[i0];
        double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        long b = head[q], e = head[q + 1];
        if (e - b > ORC_SEG_CHUNK) {
            for (long c = b; c < e; c += ORC_SEG_CHUNK) {
                double part[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                long ce = c + ORC_SEG_CHUNK < e ? c + ORC_SEG_CHUNK : e;
                for (long s = c; s < ce; ++s)
                    for (int k = 0; k < 9; ++k) part[k] += tvals[9 * idx[s] + k];
                for (int k = 0; k < 9; ++k) acc[k] += part[k];
            }
        } else {
            for (long s = b; s < e; ++s)
                for (int k = 0; k < 9; ++k) acc[k] += tvals[9 * idx[s] + k];
        }
        for (int k = 0; k < 9; ++k) ovals[9 * q + k] = acc[k];
    }
    free(head);
    free(keys);
    free(idx);
    return u;
}

/* ------------------------------------------------------------------------- */
/* Spmv::srbk_spmv -- StiffGIPC/linear_system/utils/spmv.cu:188-313            */
/* y = b*y; for each upper block (i,j,H): y_i += a H x_j; if i != j          */
/* y_j += a H^T x_i.                                                          */
/* ------------------------------------------------------------------------- */
void orc_spmv(double a, long nnz, const double* vals, const int* rows, const int* cols,
              const double* x, double b, double* y, int n_rows)
{
    long n3 = 3L * n_rows;
    for (long i = 0; i < n3; ++i) y[i] = (b != 0.0) ? b * y[i] : 0.0;
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    double* priv = (double*)calloc((size_t)nt * n3, sizeof(double));
#pragma omp parallel
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double* yp = priv + (size_t)t * n3;
#pragma omp for schedule(static)
        for (long k = 0; k < nnz; ++k) {
            int i = rows[k], j = cols[k];
            const double* H = vals + 9 * k;
            const double* xj = x + 3L * j;
            const double* xi = x + 3L * i;
            for (int r = 0; r < 3; ++r)
                yp[3L * i + r] += a * (H[0 * 3 + r] * xj[0] + H[1 * 3 + r] * xj[1] + H[2 * 3 + r] * xj[2]);
            if (i != j && g_fault != 1)
                for (int c = 0; c < 3; ++c)
                    yp[3L * j + c] += a * (H[c * 3 + 0] * xi[0] + H[c * 3 + 1] * xi[1] + H[c * 3 + 2] * xi[2]);
        }
#pragma omp for schedule(static)
        for (long e = 0; e < n3; ++e) {
            double s = 0;
            for (int q = 0; q < nt; ++q) s += priv[(size_t)q * n3 + e];
            y[e] += s;
        }
    }
    free(priv);
}

/* ------------------------------------------------------------------------- */
/* MAS preconditioner -- StiffGIPC/MASPreconditioner.cu                       */
/* Hierarchy (ReorderRealtime :1785-1826 and the kernels :29-625,1132-1548):  */
/* at level l the nodes are split into domains of B consecutive ids; the      */
/* graph of (mesh edges + contact cliques) mapped to level l and restricted   */
/* to edges inside one domain is split into connected components; component  */
/* ids at level l+1 run over the domains in order and, inside a domain, over  */
/* the components by their lowest member.  Slots of level l start at off_l,   */
/* off_1 = ceil(n/B)*B, off_{l+1} = off_l + ceil(c_l/B)*B.                    */
/* Galerkin blocks (PrepareHessian_bcoo :1829-2106) are P_l^T A P_l limited   */
/* to pairs of nodes in one level-l domain; each domain is inverted densely   */
/* in FP64 by Gauss-Jordan without pivoting, zero diagonal -> 1 (:630-726).   */
/* Apply (:2248-2289): z = sum_l P_l blockinv(A_l) P_l^T r.                   */
/* ------------------------------------------------------------------------- */
#define ORC_MAX_LEVELS 6

struct orc_mas {
    int n, B, L, coff;
    unsigned *nbr_start, *nbr_num, *nbr_list;
    long nbr_total;
    int counts[ORC_MAX_LEVELS + 1];
    int offsets[ORC_MAX_LEVELS + 2];
    int total;
    int* map;        /* map[l*n + v] local id of v at level l, l = 0..L-1 */
    int* going_next; /* [total] */
    int nblk;
    double* Mgal;    /* [domains * nblk * 9] */
    double* Minv;
    float* Minv32;
    int* cptr;       /* [total + 1]: vertices of every node of every level, */
    int* clist;      /* ascending (the restriction's gather lists)           */
};

int orc_mas_num_levels_for(int n, int B)
{
    /* computeNumLevels, MASPreconditioner.cu:1731-1750 */
    int nLevel = 1;
    int levelSz = (n + B - 1) / B * B;
    while (levelSz > B) {
        levelSz /= B;
        nLevel++;
        levelSz = (levelSz + B - 1) / B * B;
    }
    nLevel = nLevel + 1;
    return nLevel > ORC_MAX_LEVELS ? ORC_MAX_LEVELS : nLevel;
}

orc_mas* orc_mas_create(int n, int bank, const unsigned* nbr_start, const unsigned* nbr_num,
                        const unsigned* nbr_list, int collision_node_offset)
{
    if (n < 1 || (bank != 16 && bank != 32)) return NULL;
    orc_mas* m = (orc_mas*)calloc(1, sizeof(orc_mas));
    m->n = n;
    m->B = bank;
    m->coff = collision_node_offset;
    m->L = orc_mas_num_levels_for(n, bank);
    long tot = 0;
    for (int v = 0; v < n; ++v) {
        long e = (long)nbr_start[v] + nbr_num[v];
        if (e > tot) tot = e;
    }
    m->nbr_total = tot;
    m->nbr_start = (unsigned*)malloc(sizeof(unsigned) * n);
    m->nbr_num = (unsigned*)malloc(sizeof(unsigned) * n);
    m->nbr_list = (unsigned*)malloc(sizeof(unsigned) * (tot > 0 ? tot : 1));
    memcpy(m->nbr_start, nbr_start, sizeof(unsigned) * n);
    memcpy(m->nbr_num, nbr_num, sizeof(unsigned) * n);
    if (tot) memcpy(m->nbr_list, nbr_list, sizeof(unsigned) * tot);
    m->map = (int*)malloc(sizeof(int) * (size_t)n * m->L);
    m->nblk = bank * (bank + 1) / 2;
    return m;
}

void orc_mas_destroy(orc_mas* m)
{
    if (!m) return;
    free(m->nbr_start); free(m->nbr_num); free(m->nbr_list); free(m->map);
    free(m->going_next); free(m->Mgal); free(m->Minv); free(m->Minv32);
    free(m->cptr); free(m->clist);
    free(m);
}

static int uf_find(int* par, int a)
{
    while (par[a] != a) {
        par[a] = par[par[a]];
        a = par[a];
    }
    return a;
}
static void uf_union(int* par, int a, int b)
{
    a = uf_find(par, a);
    b = uf_find(par, b);
    if (a != b) {
        if (a < b) par[b] = a; else par[a] = b;
    }
}

/* Decode a contact pair into up to 4 vertex ids (minus node offset, -1 if none).
 * _buildCollisionConnection_new, MASPreconditioner.cu:1132-1548. */
static int decode_pair(const int* pr, int coff, int out[4])
{
    int v[4] = {pr[0], pr[1], pr[2], pr[3]};
    int cnt = 4;
    if (v[0] >= 0) {
        if (v[3] < 0) v[3] = -v[3] - 1;
    } else {
        v[0] = -v[0] - 1;
        if (v[2] < 0) {
            if (v[1] < 0) {
                v[1] = -v[1] - 1; v[2] = -v[2] - 1; v[3] = -v[3] - 1;
            } else {
                cnt = 2;
            }
        } else if (v[3] < 0) {
            if (v[1] < 0) {
                v[1] = -v[1] - 1; v[3] = -v[3] - 1;
            } else {
                cnt = 3;
            }
        }
    }
    for (int i = 0; i < 4; ++i) out[i] = (i < cnt) ? v[i] - coff : -1;
    return cnt;
}

static void mas_hierarchy(orc_mas* m, const int* pairs, int n_pairs)
{
    int n = m->n, B = m->B, L = m->L;
    for (int v = 0; v < n; ++v) m->map[v] = v;
    m->counts[0] = n;
    m->offsets[0] = 0;
    m->offsets[1] = (n + B - 1) / B * B;
    int* par = (int*)malloc(sizeof(int) * n);
    int* newid = (int*)malloc(sizeof(int) * n);
    for (int l = 0; l < L; ++l) {
        /* build level l+1 from level l (also for l = L-1, as the reference
         * computes c_L for the final offset, although it is never stored) */
        int cl = m->counts[l];
        const int* ml = m->map + (size_t)l * n;
        for (int a = 0; a < cl; ++a) par[a] = a;
        for (int u = 0; u < n; ++u) {
            int a = ml[u];
            for (unsigned k = 0; k < m->nbr_num[u]; ++k) {
                int v = (int)m->nbr_list[m->nbr_start[u] + k];
                int b = ml[v];
                if (a != b && a / B == b / B) uf_union(par, a, b);
            }
        }
        for (int p = 0; p < n_pairs; ++p) {
            int vid[4];
            decode_pair(pairs + 4 * (size_t)p, m->coff, vid);
            int cv[4];
            for (int i = 0; i < 4; ++i) cv[i] = (vid[i] >= 0 && vid[i] < n) ? ml[vid[i]] : -1;
            for (int i = 0; i < 4; ++i)
                for (int j = i + 1; j < 4; ++j)
                    if (cv[i] >= 0 && cv[j] >= 0 && cv[i] != cv[j] && cv[i] / B == cv[j] / B)
                        uf_union(par, cv[i], cv[j]);
        }
        /* number components domain by domain, in order of their lowest member */
        int next = 0;
        for (int a = 0; a < cl; ++a) newid[a] = -1;
        for (int a = 0; a < cl; ++a) {
            int r = uf_find(par, a);
            if (newid[r] < 0) newid[r] = next++;
            newid[a] = newid[r];
        }
        m->counts[l + 1] = next;
        m->offsets[l + 1] = m->offsets[l] + (cl + B - 1) / B * B;
        if (l + 1 < L) {
            int* mn = m->map + (size_t)(l + 1) * n;
            for (int u = 0; u < n; ++u) mn[u] = newid[ml[u]];
        }
    }
    m->total = m->offsets[L];
    free(m->going_next);
    m->going_next = (int*)malloc(sizeof(int) * (size_t)(m->total > 0 ? m->total : 1));
    for (int s = 0; s < m->total; ++s) m->going_next[s] = -1;
    for (int l = 0; l + 1 < L; ++l) {
        const int* ml = m->map + (size_t)l * n;
        const int* mn = m->map + (size_t)(l + 1) * n;
        for (int u = 0; u < n; ++u) m->going_next[m->offsets[l] + ml[u]] = m->offsets[l + 1] + mn[u];
    }
    /* the last level points into level L (never read by the apply); newid[]
     * still holds the level-L ids of the level-(L-1) nodes */
    {
        int l = L - 1;
        const int* ml = m->map + (size_t)l * n;
        for (int u = 0; u < n; ++u) m->going_next[m->offsets[l] + ml[u]] = m->offsets[L] + newid[ml[u]];
    }
    free(par);
    free(newid);
}

static inline int blk_index(int B, int r, int c) { return B * r - r * (r + 1) / 2 + c; }

static void add_block(double* dst, const double* H, int transpose, double sgn)
{
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            dst[c * 3 + r] += sgn * (transpose ? H[r * 3 + c] : H[c * 3 + r]);
}

/* Gauss-Jordan inverse without pivoting of an n x n row-major matrix. */
static void gj_inverse(int n, double* A, double* I)
{
    for (int i = 0; i < n * n; ++i) I[i] = 0.0;
    for (int i = 0; i < n; ++i) I[i * n + i] = 1.0;
    for (int j = 0; j < n; ++j) {
        double piv = A[j * n + j];
        double inv = 1.0 / piv;
        for (int c = 0; c < n; ++c) {
            A[j * n + c] *= inv;
            I[j * n + c] *= inv;
        }
        for (int k = 0; k < n; ++k) {
            if (k == j) continue;
            double f = A[k * n + j];
            if (f == 0.0) continue;
            for (int c = 0; c < n; ++c) {
                A[k * n + c] -= f * A[j * n + c];
                I[k * n + c] -= f * I[j * n + c];
            }
        }
    }
}

int orc_mas_build(orc_mas* m, long nnz, const double* vals, const int* rows, const int* cols,
                  int offset, const int* pairs, int n_pairs)
{
    int n = m->n, B = m->B, L = m->L, nb = m->nblk;
    mas_hierarchy(m, pairs, n_pairs);
    int ndom = m->total / B;
    /* gather lists of the restriction: node g = off_l + map_l[v] */
    free(m->cptr); free(m->clist);
    m->cptr = (int*)calloc((size_t)m->total + 1, sizeof(int));
    m->clist = (int*)malloc(sizeof(int) * (size_t)n * L);
    for (int l = 0; l < L; ++l)
        for (int v = 0; v < n; ++v) m->cptr[m->offsets[l] + m->map[(size_t)l * n + v] + 1]++;
    for (int g = 0; g < m->total; ++g) m->cptr[g + 1] += m->cptr[g];
    {
        int* fill = (int*)malloc(sizeof(int) * (size_t)m->total);
        memcpy(fill, m->cptr, sizeof(int) * (size_t)m->total);
        for (int l = 0; l < L; ++l)
            for (int v = 0; v < n; ++v) m->clist[fill[m->offsets[l] + m->map[(size_t)l * n + v]]++] = v;
        free(fill);
    }
    free(m->Mgal); free(m->Minv); free(m->Minv32);
    m->Mgal = (double*)calloc((size_t)ndom * nb * 9, sizeof(double));
    m->Minv = (double*)calloc((size_t)ndom * nb * 9, sizeof(double));
    m->Minv32 = (float*)calloc((size_t)ndom * nb * 9, sizeof(float));
    for (long k = 0; k < nnz; ++k) {
        int i = rows[k] - offset, j = cols[k] - offset;
        if (i < 0 || j < 0 || i >= n || j >= n) continue;
        const double* H = vals + 9 * k;
        for (int l = 0; l < L; ++l) {
            int a = m->map[(size_t)l * n + i], b = m->map[(size_t)l * n + j];
            if (a / B != b / B) continue;
            int dom = (m->offsets[l] + a) / B;
            double* D = m->Mgal + (size_t)dom * nb * 9;
            int ra = a % B, rb = b % B;
            if (a == b) {
                double* blk = D + 9 * blk_index(B, ra, ra);
                add_block(blk, H, 0, 1.0);
                if (i != j && !(g_fault == 2 && l > 0)) add_block(blk, H, 1, 1.0);
            } else if (a < b) {
                add_block(D + 9 * blk_index(B, ra, rb), H, 0, 1.0);
            } else {
                add_block(D + 9 * blk_index(B, rb, ra), H, 1, 1.0);
            }
        }
    }
    int n3 = 3 * B;
#pragma omp parallel
    {
        double* A = (double*)malloc(sizeof(double) * n3 * n3);
        double* I = (double*)malloc(sizeof(double) * n3 * n3);
#pragma omp for schedule(dynamic, 4)
        for (int d = 0; d < ndom; ++d) {
            const double* D = m->Mgal + (size_t)d * nb * 9;
            for (int rb = 0; rb < B; ++rb)
                for (int cb = 0; cb < B; ++cb)
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) {
                            double v;
                            if (cb >= rb) v = D[9 * blk_index(B, rb, cb) + j * 3 + i];
                            else v = D[9 * blk_index(B, cb, rb) + i * 3 + j];
                            A[(3 * rb + i) * n3 + 3 * cb + j] = v;
                        }
            for (int i = 0; i < n3; ++i)
                if (A[i * n3 + i] == 0.0) A[i * n3 + i] = 1.0;
            gj_inverse(n3, A, I);
            /* symmetrise the 3x3 diagonal blocks from their upper part */
            for (int rb = 0; rb < B; ++rb)
                for (int i = 0; i < 3; ++i)
                    for (int j = i + 1; j < 3; ++j)
                        I[(3 * rb + j) * n3 + 3 * rb + i] = I[(3 * rb + i) * n3 + 3 * rb + j];
            double* O = m->Minv + (size_t)d * nb * 9;
            float* O32 = m->Minv32 + (size_t)d * nb * 9;
            for (int rb = 0; rb < B; ++rb)
                for (int cb = rb; cb < B; ++cb)
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) {
                            double v = I[(3 * rb + i) * n3 + 3 * cb + j];
                            O[9 * blk_index(B, rb, cb) + j * 3 + i] = v;
                            O32[9 * blk_index(B, rb, cb) + j * 3 + i] = (float)v;
                        }
        }
        free(A);
        free(I);
    }
    return 0;
}

void orc_mas_apply(orc_mas* m, const double* r, double* z, int fp32)
{
    int n = m->n, B = m->B, L = m->L, nb = m->nblk, tot = m->total;
    int ndom = tot / B;
    if (!fp32) {
        double* R = (double*)calloc((size_t)tot * 3, sizeof(double));
        double* Z = (double*)calloc((size_t)tot * 3, sizeof(double));
#pragma omp parallel for schedule(static)
        for (int g = 0; g < tot; ++g)
            for (int e = m->cptr[g]; e < m->cptr[g + 1]; ++e)
                for (int d = 0; d < 3; ++d) R[3 * (size_t)g + d] += r[3 * (size_t)m->clist[e] + d];
#pragma omp parallel for schedule(static)
        for (int dm = 0; dm < ndom; ++dm) {
            const double* M = m->Minv + (size_t)dm * nb * 9;
            const double* Rd = R + 3 * (size_t)dm * B;
            double* Zd = Z + 3 * (size_t)dm * B;
            for (int rb = 0; rb < B; ++rb)
                for (int cb = rb; cb < B; ++cb) {
                    const double* H = M + 9 * blk_index(B, rb, cb);
                    for (int i = 0; i < 3; ++i)
                        Zd[3 * rb + i] += H[0 * 3 + i] * Rd[3 * cb] + H[1 * 3 + i] * Rd[3 * cb + 1] +
                                          H[2 * 3 + i] * Rd[3 * cb + 2];
                    if (cb != rb)
                        for (int j = 0; j < 3; ++j)
                            Zd[3 * cb + j] += H[j * 3 + 0] * Rd[3 * rb] + H[j * 3 + 1] * Rd[3 * rb + 1] +
                                              H[j * 3 + 2] * Rd[3 * rb + 2];
                }
        }
#pragma omp parallel for schedule(static)
        for (int v = 0; v < n; ++v)
            for (int d = 0; d < 3; ++d) {
                double s = 0;
                for (int l = 0; l < L; ++l) s += Z[3 * (size_t)(m->offsets[l] + m->map[(size_t)l * n + v]) + d];
                z[3 * (size_t)v + d] = s;
            }
        free(R);
        free(Z);
    } else {
        float* R = (float*)calloc((size_t)tot * 3, sizeof(float));
        float* Z = (float*)calloc((size_t)tot * 3, sizeof(float));
#pragma omp parallel for schedule(static)
        for (int g = 0; g < tot; ++g)
            for (int e = m->cptr[g]; e < m->cptr[g + 1]; ++e)
                for (int d = 0; d < 3; ++d) R[3 * (size_t)g + d] += (float)r[3 * (size_t)m->clist[e] + d];
#pragma omp parallel for schedule(static)
        for (int dm = 0; dm < ndom; ++dm) {
            const float* M = m->Minv32 + (size_t)dm * nb * 9;
            const float* Rd = R + 3 * (size_t)dm * B;
            float* Zd = Z + 3 * (size_t)dm * B;
            for (int rb = 0; rb < B; ++rb)
                for (int cb = rb; cb < B; ++cb) {
                    const float* H = M + 9 * blk_index(B, rb, cb);
                    for (int i = 0; i < 3; ++i)
                        Zd[3 * rb + i] += H[0 * 3 + i] * Rd[3 * cb] + H[1 * 3 + i] * Rd[3 * cb + 1] +
                                          H[2 * 3 + i] * Rd[3 * cb + 2];
                    if (cb != rb)
                        for (int j = 0; j < 3; ++j)
                            Zd[3 * cb + j] += H[j * 3 + 0] * Rd[3 * rb] + H[j * 3 + 1] * Rd[3 * rb + 1] +
                                              H[j * 3 + 2] * Rd[3 * rb + 2];
                }
        }
#pragma omp parallel for schedule(static)
        for (int v = 0; v < n; ++v)
            for (int d = 0; d < 3; ++d) {
                float s = 0;
                for (int l = 0; l < L; ++l) s += Z[3 * (size_t)(m->offsets[l] + m->map[(size_t)l * n + v]) + d];
                z[3 * (size_t)v + d] = (double)s;
            }
        free(R);
        free(Z);
    }
}

int orc_mas_levels(const orc_mas* m, int* level_num, int* offsets, int* counts)
{
    *level_num = m->L;
    for (int l = 0; l <= m->L; ++l) offsets[l] = m->offsets[l];
    for (int l = 0; l < m->L; ++l) counts[l] = m->counts[l];
    return m->total;
}

void orc_mas_tables(const orc_mas* m, int* coarse_table, int* going_next)
{
    int n = m->n, L = m->L;
    if (coarse_table)
        for (int v = 0; v < n; ++v)
            for (int l = 1; l < L; ++l)
                coarse_table[(size_t)v * (L - 1) + (l - 1)] = m->offsets[l] + m->map[(size_t)l * n + v];
    if (going_next) memcpy(going_next, m->going_next, sizeof(int) * (size_t)m->total);
}

void orc_mas_matrices(const orc_mas* m, double* out, int inverse)
{
    size_t cnt = (size_t)(m->total / m->B) * m->nblk * 9;
    memcpy(out, inverse ? m->Minv : m->Mgal, sizeof(double) * cnt);
}

/* ------------------------------------------------------------------------- */
/* PCGSolver::pcg -- StiffGIPC/linear_system/solver/pcg_solver.cu:191-332     */
/* ------------------------------------------------------------------------- */
/* fixed 64-way split summed in order: the result does not depend on the
 * thread count or on how OpenMP combines reductions */
static double dot(long n, const double* a, const double* b)
{
    double part[64];
    const long chunk = (n + 63) / 64;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < 64; ++c) {
        double s = 0;
        const long lo = c * chunk, hi = lo + chunk < n ? lo + chunk : n;
        for (long i = lo; i < hi; ++i) s += a[i] * b[i];
        part[c] = s;
    }
    double s = 0;
    for (int c = 0; c < 64; ++c) s += part[c];
    return s;
}

static void precond_apply(const orc_precond* pc, long n3, const double* r, double* z)
{
    memcpy(z, r, sizeof(double) * n3);
    if (!pc) return;
    if (pc->abd_inv) {
        for (int b = 0; b < pc->n_bodies; ++b) {
            const double* P = pc->abd_inv + 144 * (size_t)b;
            for (int i = 0; i < 12; ++i) {
                double sm = 0.0;
                for (int j = 0; j < 12; ++j) sm += P[i * 12 + j] * r[12 * (size_t)b + j];
                z[12 * (size_t)b + i] = sm;
            }
        }
    }
    if (pc->mas) orc_mas_apply(pc->mas, r + 3 * (size_t)pc->mas_row_offset, z + 3 * (size_t)pc->mas_row_offset,
                               pc->mas_fp32);
}

/* The full symmetric matrix as block rows (upper blocks and the transposes
 * of the off-diagonal ones, each row's blocks in column order), built once
 * per solve so that the PCG's products run row-parallel without private
 * output copies: y_i = a sum_j A_ij x_j, each row summed in column order. */
typedef struct {
    int n;
    long* ptr;
    int* col;
    double* val; /* column-major 3x3 per entry */
} orc_bsr;

static void bsr_build(orc_bsr* m, long nnz, const double* vals, const int* rows, const int* cols, int n)
{
    m->n = n;
    m->ptr = (long*)calloc((size_t)n + 1, sizeof(long));
    for (long k = 0; k < nnz; ++k) {
        m->ptr[rows[k] + 1]++;
        if (rows[k] != cols[k]) m->ptr[cols[k] + 1]++;
    }
    for (int i = 0; i < n; ++i) m->ptr[i + 1] += m->ptr[i];
    long tot = m->ptr[n];
    m->col = (int*)malloc(sizeof(int) * (size_t)(tot > 0 ? tot : 1));
    m->val = (double*)malloc(sizeof(double) * 9 * (size_t)(tot > 0 ? tot : 1));
    long* fill = (long*)malloc(sizeof(long) * ((size_t)n + 1));
    memcpy(fill, m->ptr, sizeof(long) * ((size_t)n + 1));
    /* rows sorted, upper (row <= col): for row i the transposed entries
     * (col i, from rows r < i) come first in r order, then its upper blocks in
     * column order, so every row ends up in column order */
    for (long k = 0; k < nnz; ++k) {
        int i = rows[k], j = cols[k];
        if (i != j) {
            long e = fill[j]++;
            m->col[e] = i;
            const double* H = vals + 9 * k;
            double* T = m->val + 9 * e;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) T[c * 3 + r] = H[r * 3 + c];
        }
    }
    for (long k = 0; k < nnz; ++k) {
        long e = fill[rows[k]]++;
        m->col[e] = cols[k];
        memcpy(m->val + 9 * e, vals + 9 * k, sizeof(double) * 9);
    }
    free(fill);
}

static void bsr_free(orc_bsr* m)
{
    free(m->ptr);
    free(m->col);
    free(m->val);
}

static void bsr_spmv(const orc_bsr* m, const double* x, double* y)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < m->n; ++i) {
        double a0 = 0, a1 = 0, a2 = 0;
        for (long e = m->ptr[i]; e < m->ptr[i + 1]; ++e) {
            const double* H = m->val + 9 * e;
            const double* xj = x + 3L * m->col[e];
            a0 += H[0] * xj[0] + H[3] * xj[1] + H[6] * xj[2];
            a1 += H[1] * xj[0] + H[4] * xj[1] + H[7] * xj[2];
            a2 += H[2] * xj[0] + H[5] * xj[1] + H[8] * xj[2];
        }
        y[3L * i] = a0;
        y[3L * i + 1] = a1;
        y[3L * i + 2] = a2;
    }
}

int orc_pcg2(long nnz, const double* vals, const int* rows, const int* cols, int n_rows, const orc_precond* pc,
             const double* b, double* x, double tol, int stop_mode, int max_iter, double* stats)
{
    long n3 = 3L * n_rows;
    double* r = (double*)malloc(sizeof(double) * n3);
    double* z = (double*)malloc(sizeof(double) * n3);
    double* p = (double*)malloc(sizeof(double) * n3);
    double* Ap = (double*)malloc(sizeof(double) * n3);
    orc_bsr A;
    bsr_build(&A, nnz, vals, rows, cols, n_rows);
    for (long i = 0; i < n3; ++i) { x[i] = 0.0; r[i] = b[i]; }
    double bnorm = sqrt(dot(n3, b, b));
    int k = 0, iters = 0;
    double rz = 0, rz0 = 0;
    if (bnorm > 0) {
        precond_apply(pc, n3, r, z);
        rz = dot(n3, r, z);
        memcpy(p, z, sizeof(double) * n3);
        rz0 = rz;
        for (k = 1; k < max_iter; ++k) {
            bsr_spmv(&A, p, Ap);
            double alpha = rz / dot(n3, p, Ap);
#pragma omp parallel for schedule(static)
            for (long i = 0; i < n3; ++i) {
                x[i] += alpha * p[i];
                r[i] -= alpha * Ap[i];
            }
            iters = k;
            if (stop_mode == 0) {
                if (fabs(rz) <= tol * rz0) break;
            } else {
                if (sqrt(dot(n3, r, r)) <= tol * bnorm) break;
            }
            precond_apply(pc, n3, r, z);
            double rz_new = dot(n3, r, z);
            double beta = rz_new / rz;
#pragma omp parallel for schedule(static)
            for (long i = 0; i < n3; ++i) p[i] = z[i] + beta * p[i];
            rz = rz_new;
        }
    }
    if (stats) {
        bsr_spmv(&A, x, Ap);
        double rr = 0;
        for (long i = 0; i < n3; ++i) rr += (b[i] - Ap[i]) * (b[i] - Ap[i]);
        stats[0] = iters;
        stats[1] = bnorm > 0 ? sqrt(rr) / bnorm : 0.0;
        stats[2] = rz;
        stats[3] = rz0;
    }
    free(r); free(z); free(p); free(Ap);
    bsr_free(&A);
    return k;
}

int orc_pcg(long nnz, const double* vals, const int* rows, const int* cols, int n_rows,
            orc_mas* mas, const double* b, double* x, double tol, int stop_mode, int max_iter,
            int mas_fp32, double* stats)
{
    orc_precond pc = {mas, 0, mas_fp32, 0, NULL};
    return orc_pcg2(nnz, vals, rows, cols, n_rows, mas ? &pc : NULL, b, x, tol, stop_mode, max_iter, stats);
}

/* ------------------------------------------------------------------------- */
/* Energy sums of the Newton objective (line search; GIPC.cu:7373-7870), with */
/* the scaling of the assembled gradient: dt^2 vol Psi_SNH, dt^2 area Psi_BW, */
/* 1/2 |rest edge| k_bend dt^2 x^T Q x, 1/2 m |x - x~|^2.                     */
/* ------------------------------------------------------------------------- */
double orc_energy_tets(int n, const double* x, const int* tets, const double* dm_inv, const double* vol,
                       const double* mu, const double* lam, double dt)
{
    double e = 0.0;
    for (int t = 0; t < n; ++t) {
        double xs[12];
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 3; ++d) xs[3 * a + d] = x[3 * (size_t)tets[4 * (size_t)t + a] + d];
        e += dt * dt * vol[t] * orc_tet_snh_energy(xs, dm_inv + 9 * (size_t)t, mu[t], lam[t]);
    }
    return e;
}

double orc_energy_tris(int n, const double* x, const int* tris, const double* dm_inv2, const double* area,
                       double k_st, double k_sh, double rate, double dt)
{
    double e = 0.0;
    for (int t = 0; t < n; ++t) {
        double xs[9];
        for (int a = 0; a < 3; ++a)
            for (int d = 0; d < 3; ++d) xs[3 * a + d] = x[3 * (size_t)tris[3 * (size_t)t + a] + d];
        e += dt * dt * orc_tri_bw_energy(xs, dm_inv2 + 4 * (size_t)t, area[t], k_st, k_sh, rate);
    }
    return e;
}

double orc_energy_bending(int n, const double* x, const double* rest_x, const int* edges, const int* adj,
                          const double* Q16, double k_bend, double dt)
{
    double e = 0.0;
    for (int i = 0; i < n; ++i) {
        if (adj[2 * i + 1] == -1) continue;
        int vid[4] = {edges[2 * i], edges[2 * i + 1], adj[2 * i], adj[2 * i + 1]};
        double L = 0.0;
        for (int d = 0; d < 3; ++d) {
            double r = rest_x[3 * (size_t)vid[0] + d] - rest_x[3 * (size_t)vid[1] + d];
            L += r * r;
        }
        double s = 0.0;
        for (int d = 0; d < 3; ++d)
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b)
                    s += x[3 * (size_t)vid[a] + d] * Q16[16 * (size_t)i + 4 * a + b] * x[3 * (size_t)vid[b] + d];
        e += 0.5 * sqrt(L) * k_bend * dt * dt * s;
    }
    return e;
}
