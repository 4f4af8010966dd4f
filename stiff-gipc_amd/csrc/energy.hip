// Energies of the Newton objective for the line search, on gfx950.
//
// Reference: the energy reductions GIPC::computeEnergy uses in its line
// search (StiffGIPC/GIPC.cu:7373-7870: kinetic, stable Neo-Hookean tets,
// Baraff-Witkin triangles, quadratic bending, barrier), each written here as
// one thread per element into a per-element array that a fixed-order block
// reduction sums (the reference reduces with warp shuffles in a launch-
// dependent order).  The scaling follows the assembled gradient and Hessian:
// E = 1/2 |x - x~|^2_M + dt^2 (sum vol Psi_SNH + sum area Psi_BW + sum
// 1/2 |e| k_bend x^T Q x) + sum kappa dHat^2 (I5 - 1)^2 ln(I5)^2.
#include "energy.hpp"
#include "wave_ops.hpp"

namespace sgipc {

namespace {

inline int g256(long n) { return (int)((n + 255) / 256); }

__global__ void k_e_inertia(int n, const double* __restrict__ x, const double* __restrict__ xt,
                            const double* __restrict__ mass, double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double e = x[3L * i + d] - xt[3L * i + d];
        s += e * e;
    }
    out[i] = 0.5 * mass[i] * s;
}

// stable Neo-Hookean: Psi = mu/2 (|F|^2 - 3) - mu (J - 1) + lam/2 (J - 1)^2
// with mu, lam the per-tet lengthRate / volumeRate (femEnergy.cu:1641-1717)
__global__ void k_e_tets(int n, const double* __restrict__ x, const uint32_t* __restrict__ tets,
                         const double* __restrict__ dm_inv, const double* __restrict__ vol,
                         const double* __restrict__ mu, const double* __restrict__ lam, double dt2,
                         double* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long v0 = tets[4L * t];
    double Ds[3][3];
    for (int k = 0; k < 3; ++k) {
        const long v = tets[4L * t + 1 + k];
        for (int i = 0; i < 3; ++i) Ds[i][k] = x[3 * v + i] - x[3 * v0 + i];
    }
    const double* Dm = dm_inv + 9L * t;
    double F[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) F[i][j] = Ds[i][0] * Dm[j] + Ds[i][1] * Dm[3 + j] + Ds[i][2] * Dm[6 + j];
    double I2 = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) I2 += F[i][j] * F[i][j];
    const double J = F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) - F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
                     F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
    const double m = mu[t], l = lam[t];
    out[t] = dt2 * vol[t] * (0.5 * m * (I2 - 3.0) - m * (J - 1.0) + 0.5 * l * (J - 1.0) * (J - 1.0));
}

// Baraff-Witkin stretch + shear (femEnergy.cu:2146-2215)
__global__ void k_e_tris(int n, const double* __restrict__ x, const uint32_t* __restrict__ tris,
                         const double* __restrict__ dm_inv2, const double* __restrict__ area, double k_st,
                         double k_sh, double rate, double dt2, double* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long a = tris[3L * t], b = tris[3L * t + 1], c = tris[3L * t + 2];
    const double* Dm = dm_inv2 + 4L * t;
    double F[3][2];
    for (int i = 0; i < 3; ++i) {
        const double d1 = x[3 * b + i] - x[3 * a + i], d2 = x[3 * c + i] - x[3 * a + i];
        for (int j = 0; j < 2; ++j) F[i][j] = d1 * Dm[j] + d2 * Dm[2 + j];
    }
    double E = 0.0;
    for (int j = 0; j < 2; ++j) {
        const double sig = sqrt(F[0][j] * F[0][j] + F[1][j] * F[1][j] + F[2][j] * F[2][j]);
        const double e = sig - 1.0;
        E += k_st * (e * e + (sig * sig > 1.0 ? rate * e * e * e : 0.0));
    }
    const double I6 = F[0][0] * F[0][1] + F[1][0] * F[1][1] + F[2][0] * F[2][1];
    out[t] = dt2 * area[t] * (E + k_sh * I6 * I6);
}

// quadratic bending (femEnergy.cu:2620-2754): 1/2 |rest edge| k_bend x^T Q x
__global__ void k_e_bending(int n, const double* __restrict__ x, const double* __restrict__ rest,
                            const uint32_t* __restrict__ edges, const uint32_t* __restrict__ adj,
                            const double* __restrict__ Q16, double k_bend, double dt2, double* __restrict__ out)
{
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= n) return;
    if (adj[2L * h + 1] == 0xFFFFFFFFu) {
        out[h] = 0.0;
        return;
    }
    const long v[4] = {edges[2L * h], edges[2L * h + 1], adj[2L * h], adj[2L * h + 1]};
    double L = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double e = rest[3 * v[0] + d] - rest[3 * v[1] + d];
        L += e * e;
    }
    const double* Q = Q16 + 16L * h;
    double s = 0.0;
    for (int d = 0; d < 3; ++d)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) s += x[3 * v[a] + d] * Q[4 * a + b] * x[3 * v[b] + d];
    out[h] = 0.5 * sqrt(L) * k_bend * dt2 * s;
}

struct P3 {
    double x, y, z;
};
__device__ __forceinline__ P3 ld(const double* p, long i) { return P3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ P3 sub(P3 a, P3 b) { return P3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ P3 cross(P3 a, P3 b)
{
    return P3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// barrier of every pair type: kappa dHat^2 (I5 - 1)^2 ln(I5)^2 with I5 the
// pair's squared distance over dHat (mlbvh.cu:229-258 distances); pairs with
// local geometry pair_x (12 per pair) are point-triangle pairs of that
// geometry
__global__ void k_e_barrier(int n, const double* __restrict__ x, const int4* __restrict__ pairs,
                            const double* __restrict__ pair_x, double dh, double kappa, double* __restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int4 pr = pairs[p];
    P3 P[4];
    int type;
    if (pair_x != nullptr) {
        for (int a = 0; a < 4; ++a) P[a] = ld(pair_x + 12L * p, a);
        type = 4;
    } else if (pr.x >= 0) {
        P[0] = ld(x, pr.x); P[1] = ld(x, pr.y); P[2] = ld(x, pr.z); P[3] = ld(x, pr.w < 0 ? -pr.w - 1 : pr.w);
        type = 5;
    } else {
        P[0] = ld(x, -pr.x - 1);
        P[1] = ld(x, pr.y);
        if (pr.z < 0) {
            type = 2;
        } else if (pr.w < 0) {
            P[2] = ld(x, pr.z);
            type = 3;
        } else {
            P[2] = ld(x, pr.z);
            P[3] = ld(x, pr.w);
            type = 4;
        }
    }
    double d2;
    if (type == 2) {
        const P3 d = sub(P[0], P[1]);
        d2 = dot(d, d);
    } else if (type == 3) {
        const P3 c = cross(sub(P[1], P[0]), sub(P[2], P[0])), e = sub(P[2], P[1]);
        d2 = dot(c, c) / dot(e, e);
    } else if (type == 4) {
        const P3 nv = cross(sub(P[2], P[1]), sub(P[3], P[1]));
        const double s = dot(sub(P[0], P[1]), nv);
        d2 = s * s / dot(nv, nv);
    } else {
        const P3 nv = cross(sub(P[1], P[0]), sub(P[3], P[2]));
        const double s = dot(sub(P[2], P[0]), nv);
        d2 = s * s / dot(nv, nv);
    }
    const double I5 = d2 / dh, L = log(I5);
    out[p] = kappa * dh * dh * (I5 - 1.0) * (I5 - 1.0) * L * L;
}

// fixed-order sum: per block, then one block over the block sums
__global__ __launch_bounds__(256) void k_sum(long n, const double* __restrict__ v, double* __restrict__ out)
{
    __shared__ double sm[16];
    double s = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) s += v[i];
    s = block_sum(s, sm);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

}  // namespace

double EnergySum::sum(hipStream_t s, const double* v, long n)
{
    if (n <= 0) return 0.0;
    const int gb = (int)std::min<long>(256, (n + 255) / 256);
    part_.reserve(gb + 1);
    k_sum<<<gb, 256, 0, s>>>(n, v, part_.get());
    SGIPC_LAUNCH_CHECK();
    k_sum<<<1, 256, 0, s>>>(gb, part_.get(), part_.get() + gb);
    SGIPC_LAUNCH_CHECK();
    double h = 0.0;
    SGIPC_HIP(hipMemcpyAsync(&h, part_.get() + gb, sizeof(double), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    return h;
}

double EnergySum::inertia(hipStream_t s, int n, const double* x, const double* xt, const double* mass)
{
    if (n <= 0 || xt == nullptr) return 0.0;
    buf_.reserve(n);
    k_e_inertia<<<g256(n), 256, 0, s>>>(n, x, xt, mass, buf_.get());
    SGIPC_LAUNCH_CHECK();
    return sum(s, buf_.get(), n);
}

double EnergySum::tets(hipStream_t s, int n, const double* x, const uint32_t* tets, const double* dm_inv,
                       const double* vol, const double* mu, const double* lam, double dt)
{
    if (n <= 0) return 0.0;
    buf_.reserve(n);
    k_e_tets<<<g256(n), 256, 0, s>>>(n, x, tets, dm_inv, vol, mu, lam, dt * dt, buf_.get());
    SGIPC_LAUNCH_CHECK();
    return sum(s, buf_.get(), n);
}

double EnergySum::tris(hipStream_t s, int n, const double* x, const uint32_t* tris, const double* dm_inv2,
                       const double* area, double k_st, double k_sh, double rate, double dt)
{
    if (n <= 0) return 0.0;
    buf_.reserve(n);
    k_e_tris<<<g256(n), 256, 0, s>>>(n, x, tris, dm_inv2, area, k_st, k_sh, rate, dt * dt, buf_.get());
    SGIPC_LAUNCH_CHECK();
    return sum(s, buf_.get(), n);
}

double EnergySum::bending(hipStream_t s, int n, const double* x, const double* rest, const uint32_t* edges,
                          const uint32_t* adj, const double* q16, double k_bend, double dt)
{
    if (n <= 0) return 0.0;
    buf_.reserve(n);
    k_e_bending<<<g256(n), 256, 0, s>>>(n, x, rest, edges, adj, q16, k_bend, dt * dt, buf_.get());
    SGIPC_LAUNCH_CHECK();
    return sum(s, buf_.get(), n);
}

double EnergySum::barrier(hipStream_t s, int n, const double* x, const int* pairs, const double* pair_x,
                          double dhat_sq, double kappa)
{
    if (n <= 0) return 0.0;
    buf_.reserve(n);
    k_e_barrier<<<g256(n), 256, 0, s>>>(n, x, reinterpret_cast<const int4*>(pairs), pair_x, dhat_sq, kappa,
                                        buf_.get());
    SGIPC_LAUNCH_CHECK();
    return sum(s, buf_.get(), n);
}

}  // namespace sgipc
