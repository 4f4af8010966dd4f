/*
 * CPU ORACLE -- TEST INFRASTRUCTURE ONLY (see sgipc_oracle.h).
 *
 * Lagged IPC friction of the contact pairs and of the ground, restated from
 * the reference:
 *   - friction set-up at the start of the step (_calFrictionLastH_DistAndTan,
 *     StiffGIPC/GIPC.cu:8431-8535; _calFrictionLastH_gd, :8405-8429): the
 *     normal force lambda = -2 sqrt(d) kappa db/dd (RANK 2 barrier for the
 *     pairs, the ground's own expression for the ground), the closest-point
 *     coordinates and the tangent basis of each pair
 *     (StiffGIPC/FrictionUtils.cuh:53-330);
 *   - gradient and Hessian (_calFrictionGradient :5130-5313,
 *     _calFrictionGradient_gd :5090-5128, _calFrictionHessian :1482-1845,
 *     _calFrictionHessian_gd :1400-1480) with the C1 static-friction clamp
 *     (FrictionUtils.cuh:330-500, SFCLAMPING_ORDER 1) for the pairs and the C0
 *     one for the ground; Hessian blocks in [PT+EE x 16 | PE x 9 | PP x 4] by
 *     MatIndex, lower blocks moved to (col, row) and zeroed, one diagonal block
 *     per ground vertex.
 * Every pair type is written in one form: relDX = sum_a w_a (x_a - x^t_a)
 * with the pair's vertex weights w (PT 1, b1 + b2 - 1, -b1, -b2; EE 1 - g1,
 * g1, g2 - 1, -g2; PE 1, e - 1, -e; PP 1, -1), u = B^T relDX in the tangent
 * basis B (3 x 2), gradient_a = mu lambda w_a B f1(|u|) u / |u|, Hessian_ab =
 * mu lambda w_a w_b B PD(M2) B^T.  The ground's tangent projector is
 * I - n n^T (the reference writes the xz entries for its y-up ground).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sgipc_oracle.h"

static void fsub(const double* a, const double* b, double* o)
{
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
static double fdot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void fcross(const double* a, const double* b, double* o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static void fnormalize(double* v)
{
    double n = sqrt(fdot(v, v));
    v[0] /= n;
    v[1] /= n;
    v[2] /= n;
}

/* decode a pair: vertex ids, count (2, 3, 4) and type (2 PP, 3 PE, 4 PT, 5 EE) */
static int fdecode(const int* pr, int* vid, int* type)
{
    if (pr[0] >= 0) {
        if (pr[3] < 0) return -1;
        for (int i = 0; i < 4; ++i) vid[i] = pr[i];
        *type = 5;
        return 4;
    }
    vid[0] = -pr[0] - 1;
    if (pr[1] < 0) return -1;
    vid[1] = pr[1];
    if (pr[2] < 0) {
        *type = 2;
        return 2;
    }
    vid[2] = pr[2];
    if (pr[3] < 0) {
        *type = 3;
        return 3;
    }
    vid[3] = pr[3];
    *type = 4;
    return 4;
}

/* squared distances of the reference (mlbvh.cu:229-258) */
static double fd2(int type, double P[4][3])
{
    double a[3], b[3], c[3], nv[3];
    if (type == 2) {
        fsub(P[0], P[1], a);
        return fdot(a, a);
    }
    if (type == 3) {
        fsub(P[1], P[0], a);
        fsub(P[2], P[0], b);
        fcross(a, b, c);
        fsub(P[2], P[1], a);
        return fdot(c, c) / fdot(a, a);
    }
    if (type == 4) {
        fsub(P[2], P[1], a);
        fsub(P[3], P[1], b);
        fcross(a, b, nv);
        fsub(P[0], P[1], c);
        double s = fdot(c, nv);
        return s * s / fdot(nv, nv);
    }
    fsub(P[1], P[0], a);
    fsub(P[3], P[2], b);
    fcross(a, b, nv);
    fsub(P[2], P[0], c);
    double s = fdot(c, nv);
    return s * s / fdot(nv, nv);
}

int orc_friction_prepare(int n, const double* xt, const int* pairs, double dhat_sq, double kappa, double* lambda,
                         double* coord, double* basis)
{
    for (int p = 0; p < n; ++p) {
        int vid[4], type;
        int k = fdecode(pairs + 4 * (size_t)p, vid, &type);
        if (k < 0) return -1;
        double P[4][3];
        for (int a = 0; a < k; ++a)
            for (int d = 0; d < 3; ++d) P[a][d] = xt[3 * (size_t)vid[a] + d];
        double b0[3], b1[3], c0 = 0.0, c1 = 0.0;
        if (type == 4) {
            /* FrictionUtils.cuh:53-86: basis (v12, (v12 x v13) x v12), beta = (B^T B)^-1 B^T v10 */
            double e1[3], e2[3], w[3], t[3];
            fsub(P[2], P[1], e1);
            fsub(P[3], P[1], e2);
            fsub(P[0], P[1], w);
            double A = fdot(e1, e1), Bm = fdot(e1, e2), C = fdot(e2, e2);
            double r0 = fdot(e1, w), r1 = fdot(e2, w), det = A * C - Bm * Bm;
            c0 = (C * r0 - Bm * r1) / det;
            c1 = (A * r1 - Bm * r0) / det;
            memcpy(b0, e1, sizeof b0);
            fnormalize(b0);
            fcross(e1, e2, t);
            fcross(t, e1, b1);
            fnormalize(b1);
        } else if (type == 5) {
            /* :157-196: basis (v01, (v01 x v23) x v01), closest-point gammas */
            double e20[3], e01[3], e23[3], t[3];
            fsub(P[0], P[2], e20);
            fsub(P[1], P[0], e01);
            fsub(P[3], P[2], e23);
            double m00 = fdot(e01, e01), m01 = -fdot(e23, e01), m11 = fdot(e23, e23);
            double r0 = -fdot(e20, e01), r1 = fdot(e20, e23), det = m00 * m11 - m01 * m01;
            c0 = (m11 * r0 - m01 * r1) / det;
            c1 = (m00 * r1 - m01 * r0) / det;
            memcpy(b0, e01, sizeof b0);
            fnormalize(b0);
            fcross(e01, e23, t);
            fcross(t, e01, b1);
            fnormalize(b1);
        } else if (type == 3) {
            /* :251-272: basis (v12, v12 x v10), eta */
            double e12[3], w[3];
            fsub(P[2], P[1], e12);
            fsub(P[0], P[1], w);
            c0 = fdot(w, e12) / fdot(e12, e12);
            memcpy(b0, e12, sizeof b0);
            fnormalize(b0);
            fcross(e12, w, b1);
            fnormalize(b1);
        } else {
            /* :301-331: the larger of the two axis crosses of v01 */
            double v01[3], xc[3], yc[3];
            fsub(P[1], P[0], v01);
            xc[0] = 0.0; xc[1] = -v01[2]; xc[2] = v01[1];
            yc[0] = v01[2]; yc[1] = 0.0; yc[2] = -v01[0];
            const double* ch = fdot(xc, xc) > fdot(yc, yc) ? xc : yc;
            memcpy(b0, ch, sizeof b0);
            fnormalize(b0);
            fcross(v01, ch, b1);
            fnormalize(b1);
        }
        coord[2 * (size_t)p] = c0;
        coord[2 * (size_t)p + 1] = c1;
        for (int d = 0; d < 3; ++d) {
            basis[6 * (size_t)p + d] = b0[d];
            basis[6 * (size_t)p + 3 + d] = b1[d];
        }
        /* GIPC.cu:8519-8526 (RANK 2) */
        double d = fd2(type, P), L = log(d / dhat_sq);
        lambda[p] = -kappa * 2.0 * sqrt(d) * (L * L * (2.0 * d - 2.0 * dhat_sq) + 2.0 * L * (d - dhat_sq) * (d - dhat_sq) / d);
    }
    return 0;
}

/* vertex weights of relDX (FrictionUtils.cuh computeRelDX_* / computeT_*) */
static void fweights(int type, const double* c, double* w)
{
    if (type == 4) {
        w[0] = 1.0; w[1] = -1.0 + c[0] + c[1]; w[2] = -c[0]; w[3] = -c[1];
    } else if (type == 5) {
        w[0] = 1.0 - c[0]; w[1] = c[0]; w[2] = c[1] - 1.0; w[3] = -c[1];
    } else if (type == 3) {
        w[0] = 1.0; w[1] = c[0] - 1.0; w[2] = -c[0];
    } else {
        w[0] = 1.0; w[1] = -1.0;
    }
}

/* u = B^T relDX for one pair; returns the vertex count */
static int frel(const double* x, const double* xt, const int* pr, const double* c, const double* B, int* vid,
                double* w, double* u, int* type)
{
    int k = fdecode(pr, vid, type);
    if (k < 0) return -1;
    fweights(*type, c, w);
    double r[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < k; ++a)
        for (int d = 0; d < 3; ++d) r[d] += w[a] * (x[3 * (size_t)vid[a] + d] - xt[3 * (size_t)vid[a] + d]);
    u[0] = fdot(B, r);
    u[1] = fdot(B + 3, r);
    return k;
}

/* symmetric 2x2 eigen clamp (the reference's makePDGeneral<double, 2>) */
static void pd2(double M[2][2])
{
    double a = M[0][0], b = 0.5 * (M[0][1] + M[1][0]), d = M[1][1];
    double tr = 0.5 * (a + d), df = 0.5 * (a - d), rad = sqrt(df * df + b * b);
    double l0 = tr + rad, l1 = tr - rad;
    double v0[2];
    if (rad > 0.0) {
        /* eigenvector of l0 */
        if (df >= 0) { v0[0] = df + rad; v0[1] = b; } else { v0[0] = b; v0[1] = rad - df; }
        double nn = sqrt(v0[0] * v0[0] + v0[1] * v0[1]);
        v0[0] /= nn;
        v0[1] /= nn;
    } else {
        v0[0] = 1.0;
        v0[1] = 0.0;
    }
    double v1[2] = {-v0[1], v0[0]};
    if (l0 < 0.0) l0 = 0.0;
    if (l1 < 0.0) l1 = 0.0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) M[i][j] = l0 * v0[i] * v0[j] + l1 * v1[i] * v1[j];
}

int orc_friction(int n, const double* x, const double* xt, const int* pairs, const int* mat_index, int n4, int n3,
                 const double* lambda, const double* coord, const double* basis, double eps2, double mu, int hess_offset,
                 double* grad, double* tvals, int* trows, int* tcols)
{
    const double eps = sqrt(eps2);
    for (int p = 0; p < n; ++p) {
        int vid[4], type;
        double w[4], u[2];
        const double* B = basis + 6 * (size_t)p;
        int k = frel(x, xt, pairs + 4 * (size_t)p, coord + 2 * (size_t)p, B, vid, w, u, &type);
        if (k < 0) return -1;
        double uu = u[0] * u[0] + u[1] * u[1], un = sqrt(uu);
        double s = mu * lambda[p];
        /* gradient: f1(|u|) u / |u| */
        double gscale = uu > eps2 ? 1.0 / un : (-un + 2.0 * eps) / (eps * eps);
        if (grad) {
            double g3[3];
            for (int d = 0; d < 3; ++d) g3[d] = s * gscale * (B[d] * u[0] + B[3 + d] * u[1]);
            for (int a = 0; a < k; ++a)
                for (int d = 0; d < 3; ++d) grad[3 * (size_t)vid[a] + d] += w[a] * g3[d];
        }
        double M[2][2];
        if (uu > eps2) {
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j) M[i][j] = (i == j ? 1.0 / un : 0.0) - u[i] * u[j] / (uu * un);
        } else {
            double f1d = (-un + 2.0 * eps) / (eps * eps), f2 = 2.0 * (eps - un) / (eps * eps);
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j)
                    M[i][j] = (i == j ? f1d : 0.0) - ((f2 != f1d && uu != 0.0) ? (f1d - f2) * u[i] * u[j] / uu : 0.0);
        }
        pd2(M);
        /* B M B^T (3 x 3) */
        double H3[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double acc = 0.0;
                for (int i = 0; i < 2; ++i)
                    for (int j = 0; j < 2; ++j) acc += B[3 * i + r] * M[i][j] * B[3 * j + c];
                H3[r][c] = s * acc;
            }
        long base;
        int mi = mat_index[p];
        if (k == 4) base = 16L * mi;
        else if (k == 3) base = 16L * n4 + 9L * mi;
        else base = 16L * n4 + 9L * n3 + 4L * mi;
        for (int ii = 0; ii < k; ++ii)
            for (int jj = 0; jj < k; ++jj) {
                long kk = base + ii * k + jj;
                int row = vid[ii] + hess_offset, col = vid[jj] + hess_offset;
                double* v = tvals + 9 * kk;
                if (row <= col) {
                    trows[kk] = row;
                    tcols[kk] = col;
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) v[c * 3 + r] = w[ii] * w[jj] * H3[r][c];
                } else {
                    trows[kk] = col;
                    tcols[kk] = row;
                    for (int e = 0; e < 9; ++e) v[e] = 0.0;
                }
            }
    }
    return 0;
}

double orc_friction_energy(int n, const double* x, const double* xt, const int* pairs, const double* lambda,
                           const double* coord, const double* basis, double eps2, double mu)
{
    const double eps = sqrt(eps2);
    double e = 0.0;
    for (int p = 0; p < n; ++p) {
        int vid[4], type;
        double w[4], u[2];
        if (frel(x, xt, pairs + 4 * (size_t)p, coord + 2 * (size_t)p, basis + 6 * (size_t)p, vid, w, u, &type) < 0)
            continue;
        double uu = u[0] * u[0] + u[1] * u[1], un = sqrt(uu);
        /* f0 of the C1 clamp (FrictionUtils.cuh:381-384) */
        double f0 = uu > eps2 ? un : uu * (-un / 3.0 + eps) / (eps * eps) + eps / 3.0;
        e += mu * lambda[p] * f0;
    }
    return e;
}

/* ground: GIPC.cu:8405-8429 */
int orc_ground_friction_prepare(int n, const double* xt, const int* verts, const double* normal, double offset,
                                double dhat_sq, double kappa, double* lambda)
{
    for (int i = 0; i < n; ++i) {
        double dist = fdot(normal, xt + 3 * (size_t)verts[i]) - offset;
        double d2 = dist * dist, t = d2 - dhat_sq;
        double gb = t * log(d2 / dhat_sq) * -2.0 - (t * t) / d2;
        lambda[i] = -kappa * 2.0 * sqrt(d2) * gb;
    }
    return 0;
}

static void ground_v(const double* x, const double* xt, int v, const double* nrm, double* V)
{
    double dx[3];
    fsub(x + 3 * (size_t)v, xt + 3 * (size_t)v, dx);
    double dn = fdot(dx, nrm);
    for (int d = 0; d < 3; ++d) V[d] = dx[d] - nrm[d] * dn;
}

int orc_ground_friction(int n, const double* x, const double* xt, const int* verts, const double* normal,
                        const double* lambda, double eps2, double mu, int hess_offset, double* grad, double* tvals,
                        int* trows, int* tcols)
{
    const double eps = sqrt(eps2);
    for (int i = 0; i < n; ++i) {
        double V[3];
        int v = verts[i];
        ground_v(x, xt, v, normal, V);
        double vv = fdot(V, V), vn = sqrt(vv);
        double m = mu * lambda[i];
        if (grad)
            for (int d = 0; d < 3; ++d) grad[3 * (size_t)v + d] += V[d] * m / (vv > eps2 ? vn : eps);
        double H[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double P = (r == c ? 1.0 : 0.0) - normal[r] * normal[c];
                H[r][c] = vv > eps2 ? m * (P / vn - V[r] * V[c] / (vv * vn)) : m / eps * P;
            }
        double* t = tvals + 9 * (size_t)i;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) t[c * 3 + r] = H[r][c];
        trows[i] = v + hess_offset;
        tcols[i] = v + hess_offset;
    }
    return 0;
}

double orc_ground_friction_energy(int n, const double* x, const double* xt, const int* verts, const double* normal,
                                  const double* lambda, double eps2, double mu)
{
    const double eps = sqrt(eps2);
    double e = 0.0;
    for (int i = 0; i < n; ++i) {
        double V[3];
        ground_v(x, xt, verts[i], normal, V);
        double vv = fdot(V, V), vn = sqrt(vv);
        /* the C0 clamp (FrictionUtils.cuh:336-339) */
        e += mu * lambda[i] * (vv > eps2 ? vn : vv / (2.0 * eps) + eps / 2.0);
    }
    return e;
}
