// Cloth assembly for gfx950: Baraff-Witkin stretch/shear triangles and
// quadratic (constant-Q) bending hinges, written as 3x3 upper triplets.
//
//  * triangle (StiffGIPC/femEnergy.cu:2146-2215; dE/dF :351-395; projected
//    Hessians :535-700): F = Ds Dm^-1 is 3x2 with columns f_u, f_v.  Each
//    stretch term psi(|f|) = (|f|-1)^2 + [|f|>1] rate (|f|-1)^3 is projected
//    onto psi'' f^f^T + max(psi'/|f|, 0)(I - f^f^T).  The shear term I6^2
//    (I6 = f_u.f_v) has Hessian 2(g g^T + I6 S) with g = (f_v, f_u) and S the
//    column swap; its spectrum is lambda+- = (I2 +- sqrt(I2^2 + 12 I6^2))/2
//    on span(g+, g-) (the +-1 eigenspace parts of g) and +-I6 on the rest of
//    the two eigenspaces of S, so the projection keeps
//    2(lambda+ q q^T + |I6| P), q ~ (lambda+ + I6) g+ + (lambda+ - I6) g-,
//    P the projector onto the sign(I6) eigenspace of S minus g_sign.  The 9x9
//    vertex Hessian is dt^2 area dF/dx^T Hq dF/dx, 6 upper blocks.
//  * hinge (femEnergy.cu:2490-2618 Q, 2620-2754 kernel): Q = 3/(A0+A1) K K^T
//    from cotangent weights of the rest hinge; gradient coef Q x and blocks
//    coef Q_ab I with coef = |rest edge| k dt^2; boundary hinges write ten
//    zero blocks at (0, 0).
#include "assembly.hpp"
#include "wave_ops.hpp"

namespace sgipc {

namespace {

__global__ __launch_bounds__(128) void k_tri_bw(int n_tris, const double* __restrict__ x,
                                                const uint3* __restrict__ tris, const double* __restrict__ dm_inv2,
                                                const double* __restrict__ area, double k_st, double k_sh,
                                                double rate, double dt2, int hess_offset, double* __restrict__ grad,
                                                double* __restrict__ tvals, int* __restrict__ trows,
                                                int* __restrict__ tcols)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tris) return;
    const uint3 tv = tris[t];
    const unsigned vid[3] = {tv.x, tv.y, tv.z};
    double xs[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) xs[a][d] = x[3 * (long)vid[a] + d];
    const double m00 = dm_inv2[4 * (long)t], m01 = dm_inv2[4 * (long)t + 1];
    const double m10 = dm_inv2[4 * (long)t + 2], m11 = dm_inv2[4 * (long)t + 3];
    double F[3][2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double d1 = xs[1][i] - xs[0][i], d2 = xs[2][i] - xs[0][i];
        F[i][0] = d1 * m00 + d2 * m10;
        F[i][1] = d1 * m01 + d2 * m11;
    }
    // D[a][j] = dF(:,j)/dx_a
    const double D[3][2] = {{-(m00 + m10), -(m01 + m11)}, {m00, m01}, {m10, m11}};
    const double scale = dt2 * area[t];

    // Hq: projected d2E/dF2 over column-major vec(F), kept as a full 6x6
    double Hq[6][6];
    double P[3][2];
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int b = 0; b < 6; ++b) Hq[a][b] = 0.0;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const double f0 = F[0][c], f1 = F[1][c], f2 = F[2][c];
        const double sig = sqrt(f0 * f0 + f1 * f1 + f2 * f2);
        const double e = sig - 1.0;
        const bool st = sig > 1.0;
        const double d1 = 2.0 * e + (st ? 3.0 * rate * e * e : 0.0);
        const double d2 = 2.0 + (st ? 6.0 * rate * e : 0.0);
        const double inv = 1.0 / sig;
        const double fh[3] = {f0 * inv, f1 * inv, f2 * inv};
        const double iso = d1 * inv > 0.0 ? d1 * inv : 0.0;
#pragma unroll
        for (int r = 0; r < 3; ++r) P[r][c] = k_st * d1 * fh[r];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
                Hq[3 * c + a][3 * c + b] = k_st * ((a == b ? iso : 0.0) + (d2 - iso) * fh[a] * fh[b]);
    }
    {
        const double a0 = F[0][0], a1 = F[1][0], a2 = F[2][0];  // f_u
        const double b0 = F[0][1], b1 = F[1][1], b2 = F[2][1];  // f_v
        const double I6 = a0 * b0 + a1 * b1 + a2 * b2;
        const double I2 = a0 * a0 + a1 * a1 + a2 * a2 + b0 * b0 + b1 * b1 + b2 * b2;
        P[0][0] += k_sh * 2.0 * I6 * b0;
        P[1][0] += k_sh * 2.0 * I6 * b1;
        P[2][0] += k_sh * 2.0 * I6 * b2;
        P[0][1] += k_sh * 2.0 * I6 * a0;
        P[1][1] += k_sh * 2.0 * I6 * a1;
        P[2][1] += k_sh * 2.0 * I6 * a2;
        const double lp = 0.5 * (I2 + sqrt(I2 * I2 + 12.0 * I6 * I6));
        // g = (f_v; f_u) = (s; s) + (d; -d), s = (f_u + f_v)/2, d = (f_v - f_u)/2
        const double s3[3] = {0.5 * (a0 + b0), 0.5 * (a1 + b1), 0.5 * (a2 + b2)};
        const double d3[3] = {0.5 * (b0 - a0), 0.5 * (b1 - a1), 0.5 * (b2 - a2)};
        const double cp = lp + I6, cm = lp - I6;
        double q[6];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            q[r] = cp * s3[r] + cm * d3[r];
            q[3 + r] = cp * s3[r] - cm * d3[r];
        }
        double qn = 0.0;
#pragma unroll
        for (int r = 0; r < 6; ++r) qn += q[r] * q[r];
        const double wq = qn > 0.0 ? 2.0 * k_sh * lp / qn : 0.0;
        // sign(I6) eigenspace of S minus its g part: 1/2 (I +- S) - h h^T, h = g_sign/|g_sign|
        const double sg = I6 >= 0.0 ? 1.0 : -1.0;
        double h[6];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            h[r] = I6 >= 0.0 ? s3[r] : d3[r];
            h[3 + r] = I6 >= 0.0 ? s3[r] : -d3[r];
        }
        double hn = 0.0;
#pragma unroll
        for (int r = 0; r < 6; ++r) hn += h[r] * h[r];
        const double wi = 2.0 * k_sh * fabs(I6);
        const double wh = hn > 0.0 ? wi / hn : 0.0;
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
            for (int b = 0; b < 6; ++b) {
                const double pb = 0.5 * ((a == b ? 1.0 : 0.0) + ((a % 3 == b % 3 && a / 3 != b / 3) ? sg : 0.0));
                Hq[a][b] += wq * q[a] * q[b] + wi * pb - wh * h[a] * h[b];
            }
    }
    if (grad) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                atomic_add(grad + 3 * (long)vid[a] + d, scale * (D[a][0] * P[d][0] + D[a][1] * P[d][1]));
    }
    // 6 upper blocks
    const long base = 6 * (long)t;
    int kk = 0;
#pragma unroll
    for (int ii = 0; ii < 3; ++ii)
#pragma unroll
        for (int jj = ii; jj < 3; ++jj, ++kk) {
            const int row = (int)vid[ii] + hess_offset, col = (int)vid[jj] + hess_offset;
            const bool keep = row <= col;
            // stored block (min, max): M(r,c) = H9(3A+r, 3B+c) with (A, B) = keep ? (ii, jj) : (jj, ii)
            const double DA[2] = {keep ? D[ii][0] : D[jj][0], keep ? D[ii][1] : D[jj][1]};
            const double DB[2] = {keep ? D[jj][0] : D[ii][0], keep ? D[jj][1] : D[ii][1]};
            trows[base + kk] = keep ? row : col;
            tcols[base + kk] = keep ? col : row;
            double* dst = tvals + 9 * (base + kk);
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    double v = 0.0;
#pragma unroll
                    for (int j = 0; j < 2; ++j)
#pragma unroll
                        for (int jp = 0; jp < 2; ++jp) v += DA[j] * DB[jp] * Hq[3 * j + r][3 * jp + c];
                    dst[c * 3 + r] = scale * v;
                }
        }
}

__device__ __forceinline__ double cot3(const double* a, const double* b)
{
    const double c0 = a[1] * b[2] - a[2] * b[1], c1 = a[2] * b[0] - a[0] * b[2], c2 = a[0] * b[1] - a[1] * b[0];
    const double cn = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    if (cn < 1e-12) return 0.0;
    return (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) / cn;
}

__global__ __launch_bounds__(256) void k_quad_bending_q(int n_edges, const double* __restrict__ rest,
                                                        const uint2* __restrict__ edges,
                                                        const uint2* __restrict__ adj, double* __restrict__ Q16)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_edges) return;
    double* Q = Q16 + 16 * (long)i;
    const uint2 e = edges[i], o = adj[i];
    double out[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) out[k] = 0.0;
    if (o.y != 0xFFFFFFFFu) {
        const unsigned vid[4] = {e.x, e.y, o.x, o.y};
        double v[4][3];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) v[a][d] = rest[3 * (long)vid[a] + d];
        double e0[3], e1[3], e2[3], e3[3], e4[3], m0[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            e0[d] = v[1][d] - v[0][d];
            e1[d] = v[2][d] - v[0][d];
            e2[d] = v[3][d] - v[0][d];
            e3[d] = v[2][d] - v[1][d];
            e4[d] = v[3][d] - v[1][d];
            m0[d] = -e0[d];
        }
        const double c1x = e0[1] * e1[2] - e0[2] * e1[1], c1y = e0[2] * e1[0] - e0[0] * e1[2],
                     c1z = e0[0] * e1[1] - e0[1] * e1[0];
        const double c2x = e0[1] * e2[2] - e0[2] * e2[1], c2y = e0[2] * e2[0] - e0[0] * e2[2],
                     c2z = e0[0] * e2[1] - e0[1] * e2[0];
        const double n1 = sqrt(c1x * c1x + c1y * c1y + c1z * c1z), n2 = sqrt(c2x * c2x + c2y * c2y + c2z * c2z);
        const double A = 0.5 * n1 + 0.5 * n2;
        if (n1 >= 1e-12 && n2 >= 1e-12 && A >= 1e-12) {
            const double k01 = cot3(e0, e1), k02 = cot3(e0, e2), k03 = cot3(m0, e3), k04 = cot3(m0, e4);
            const double K[4] = {k03 + k04, k01 + k02, -k01 - k03, -k02 - k04};
            bool ok = true;
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    out[r * 4 + c] = 3.0 / A * K[r] * K[c];
                    ok = ok && isfinite(out[r * 4 + c]);
                }
            if (!ok)
#pragma unroll
                for (int k = 0; k < 16; ++k) out[k] = 0.0;
        }
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) Q[k] = out[k];
}

__global__ __launch_bounds__(256) void k_quad_bending(int n_edges, const double* __restrict__ x,
                                                      const double* __restrict__ rest,
                                                      const uint2* __restrict__ edges, const uint2* __restrict__ adj,
                                                      const double* __restrict__ Q16, double k_bend_dt2,
                                                      int hess_offset, double* __restrict__ grad,
                                                      double* __restrict__ tvals, int* __restrict__ trows,
                                                      int* __restrict__ tcols)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_edges) return;
    const uint2 e = edges[i], o = adj[i];
    const long base = 10 * (long)i;
    if (o.y == 0xFFFFFFFFu) {
        for (int k = 0; k < 10; ++k) {
            trows[base + k] = 0;
            tcols[base + k] = 0;
#pragma unroll
            for (int c = 0; c < 9; ++c) tvals[9 * (base + k) + c] = 0.0;
        }
        return;
    }
    const unsigned vid[4] = {e.x, e.y, o.x, o.y};
    double Q[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) Q[k] = Q16[16 * (long)i + k];
    double L = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double t = rest[3 * (long)vid[0] + d] - rest[3 * (long)vid[1] + d];
        L += t * t;
    }
    const double coef = sqrt(L) * k_bend_dt2;
    if (grad) {
        double xs[4][3];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) xs[a][d] = x[3 * (long)vid[a] + d];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double g = 0.0;
#pragma unroll
                for (int b = 0; b < 4; ++b) g += Q[a * 4 + b] * xs[b][d];
                atomic_add(grad + 3 * (long)vid[a] + d, coef * g);
            }
    }
    int kk = 0;
#pragma unroll
    for (int ii = 0; ii < 4; ++ii)
#pragma unroll
        for (int jj = ii; jj < 4; ++jj, ++kk) {
            const int row = (int)vid[ii] + hess_offset, col = (int)vid[jj] + hess_offset;
            const bool keep = row <= col;
            trows[base + kk] = keep ? row : col;
            tcols[base + kk] = keep ? col : row;
            const double v = coef * (keep ? Q[ii * 4 + jj] : Q[jj * 4 + ii]);
            double* dst = tvals + 9 * (base + kk);
#pragma unroll
            for (int c = 0; c < 9; ++c) dst[c] = (c == 0 || c == 4 || c == 8) ? v : 0.0;
        }
}

}  // namespace

void launch_tri_bw(hipStream_t s, int n_tris, const double* x, const uint32_t* tris, const double* dm_inv2,
                   const double* area, double k_stretch, double k_shear, double strain_rate, double dt,
                   int hess_offset, double* grad, double* tvals, int* trows, int* tcols)
{
    if (n_tris <= 0) return;
    k_tri_bw<<<(n_tris + 127) / 128, 128, 0, s>>>(n_tris, x, reinterpret_cast<const uint3*>(tris), dm_inv2, area,
                                                  k_stretch, k_shear, strain_rate, dt * dt, hess_offset, grad, tvals,
                                                  trows, tcols);
    SGIPC_LAUNCH_CHECK();
}

void launch_quad_bending_q(hipStream_t s, int n_edges, const double* rest, const uint32_t* edges,
                           const uint32_t* adj, double* Q16)
{
    if (n_edges <= 0) return;
    k_quad_bending_q<<<(n_edges + 255) / 256, 256, 0, s>>>(n_edges, rest, reinterpret_cast<const uint2*>(edges),
                                                           reinterpret_cast<const uint2*>(adj), Q16);
    SGIPC_LAUNCH_CHECK();
}

void launch_quad_bending(hipStream_t s, int n_edges, const double* x, const double* rest, const uint32_t* edges,
                         const uint32_t* adj, const double* Q16, double k_bend, double dt, int hess_offset,
                         double* grad, double* tvals, int* trows, int* tcols)
{
    if (n_edges <= 0) return;
    k_quad_bending<<<(n_edges + 255) / 256, 256, 0, s>>>(n_edges, x, rest, reinterpret_cast<const uint2*>(edges),
                                                         reinterpret_cast<const uint2*>(adj), Q16, k_bend * dt * dt,
                                                         hess_offset, grad, tvals, trows, tcols);
    SGIPC_LAUNCH_CHECK();
}

}  // namespace sgipc
