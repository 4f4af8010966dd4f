// Assembly entry points and the whole-linear-system object of the C ABI
// (GlobalLinearSystem::solve_linear_system, StiffGIPC/linear_system/
// linear_system/global_linear_system.cu:8-148): element Hessians into the
// triplet arena in the reference's order (contacts, tets, bending hinges, cloth
// triangles, mass; StiffGIPC/GIPC.cu:10291-10485), conversion to unique upper
// BCOO, MAS rebuild with the current contact pairs, PCG.
#include <algorithm>
#include <chrono>
#include <limits>
#include <cmath>
#include <exception>
#include <memory>
#include <thread>
#include <vector>

#include "capi_internal.hpp"
#include "energy.hpp"
#include "friction.hpp"
#include "slots.hpp"

using namespace sgipc;

namespace {

template <typename F>
int run(sgipc_ctx* ctx, F&& f)
{
    return guarded(ctx, std::forward<F>(f));
}

Converter& converter_of(sgipc_ctx* ctx)
{
    if (!ctx->conv) ctx->conv = std::make_unique<Converter>();
    return *ctx->conv;
}

}  // namespace

struct sgipc_system {
    sgipc_ctx* ctx = nullptr;
    sgipc_scene_desc d{};
    int B = 32;
    DevBuf<double> hinge_q;  // Q per hinge, computed from the rest shape at create
    sgipc_mas* mas = nullptr;
    DevBuf<double> tvals;
    DevBuf<int> trows, tcols;
    DevBuf<double> bvals;
    DevBuf<int> brows, bcols;
    int nnz = 0;
    long n_trip = 0;        // triplets of the fixed (non-contact) sections
    long arena_cap = 0;     // triplets the arena holds
    int capacity = 0;       // BCOO capacity
    int n_sys = 0;          // system block rows: 4 per body + FEM vertices
    Converter conv;
    // affine bodies: point-level contact scratch, class mapping, block-Jacobi
    DevBuf<double> cvals;
    DevBuf<int> crows, ccols;
    DevBuf<double> gpt;     // point-space contact gradient (body points, then FEM points)
    ContactMapper cmap;
    AbdJacobi abdj;
    // contact detection on every solve: detected pairs, their MatIndex and
    // the capacity held
    ContactDetector det;
    DevBuf<int> dpairs, dmat, dbad;
    int dcap = 0;
    int n_detected = 0;  // pairs of the last solve
    // line search: energy terms, the pairs at trial positions, the trial
    // positions and the step reductions
    EnergySum esum;
    DevBuf<int> lpairs, lmat;
    int lcap = 0;
    DevBuf<double> stepvals;
    StepReducer stepred;
    SolveWorkspace* ws = nullptr;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    // learned coarse level: network, mode (0 off, 1 on, 2 automatic), and
    // the last iteration counts with and without it (automatic mode)
    sgipc_mpnn* net = nullptr;
    int learned_mode = 0;
    int plain_it = -1, learned_it = -1;
    long n_solves = 0;
    long next_probe = 8;     // solve index of the next measurement of the losing variant
    long probe_gap = 8;      // doubles after every probe the learned level loses
    // CEMAS partition: the system is solved in slot space (n_slots rows,
    // vertex i at row slot_of[i]); n_slots = 0 without a partition
    int n_slots = 0;
    DevBuf<int> slot_of, spairs;
    DevBuf<double> bs, xs;
    // mixed precision that failed on this system: later solves run FP64
    // until solve index fp32_probe, where FP32 is tried again (the gap
    // doubling on every further failure)
    bool fp32_failed = false;
    long fp32_probe = 0;
    long fp32_gap = 8;
};

// sgipc_mas_partition_maps is defined in partition.cpp
extern "C" int sgipc_mas_partition_maps(const int*, int, int, int*, int*, int*);

namespace {

// Slot maps of the descriptor's partition (empty parts dropped: they hold no
// vertex and give no level-1 node, so the MAS hierarchy is the same), the
// slot-space neighbour lists for the MAS and the slot-space vectors.  Every
// empty slot is linked to lane 0 of its bank (a real vertex, since the bank's
// part is not empty) and lane 0 to them: an empty slot then joins lane 0's
// level-1 node instead of forming one, and, holding no blocks and a zero
// residual, changes neither that node's Galerkin block nor its sums.  Empty
// slots follow the real lanes of their bank, so no node's lowest member and
// hence no node numbering changes.
void init_slots(sgipc_system* sys)
{
    const auto& d = sys->d;
    const int n = d.n_verts, B = sys->B;
    std::vector<int> ids(d.part_id, d.part_id + n);
    for (int v : ids) SGIPC_CHECK(v >= 0, SGIPC_BAD_INDEX, "negative part id");
    std::vector<int> used(ids);
    std::sort(used.begin(), used.end());
    used.erase(std::unique(used.begin(), used.end()), used.end());
    for (int& v : ids) v = (int)(std::lower_bound(used.begin(), used.end(), v) - used.begin());
    int map_size = 0;
    int st = sgipc_mas_partition_maps(ids.data(), n, B, &map_size, nullptr, nullptr);
    SGIPC_CHECK(st != SGIPC_CAPACITY, SGIPC_CAPACITY, "a part holds more vertices than the bank size");
    SGIPC_CHECK(st == SGIPC_OK, st, "invalid partition");
    std::vector<int> p2r(map_size), r2p(n);
    st = sgipc_mas_partition_maps(ids.data(), n, B, &map_size, p2r.data(), r2p.data());
    SGIPC_CHECK(st == SGIPC_OK, st, "invalid partition");
    hipStream_t s = sys->ctx->c.stream;
    std::vector<uint32_t> hs(n), hn(n), hl(d.total_neighbor_num > 0 ? d.total_neighbor_num : 0);
    SGIPC_HIP(hipMemcpyAsync(hs.data(), d.nbr_start, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipMemcpyAsync(hn.data(), d.nbr_num, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
    if (!hl.empty())
        SGIPC_HIP(hipMemcpyAsync(hl.data(), d.nbr_list, sizeof(uint32_t) * hl.size(), hipMemcpyDeviceToHost, s));
    SGIPC_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> ss(map_size), sn(map_size), sl;
    sl.reserve(hl.size() + 2L * map_size);
    for (int t = 0; t < map_size; ++t) {
        ss[t] = (uint32_t)sl.size();
        const int v = p2r[t], b0 = t - t % B;
        if (v >= 0) {
            for (uint32_t k = 0; k < hn[v]; ++k) {
                SGIPC_CHECK((size_t)hs[v] + k < hl.size(), SGIPC_BAD_INDEX, "neighbour list beyond its length");
                const uint32_t j = hl[hs[v] + k];
                SGIPC_CHECK(j < (uint32_t)n, SGIPC_BAD_INDEX, "neighbour id beyond the vertex count");
                sl.push_back((uint32_t)r2p[j]);
            }
            if (t == b0)
                for (int u = b0 + 1; u < b0 + B; ++u)
                    if (p2r[u] < 0) sl.push_back((uint32_t)u);
        } else {
            sl.push_back((uint32_t)b0);
        }
        sn[t] = (uint32_t)sl.size() - ss[t];
    }
    const int total = (int)sl.size();
    if (sl.empty()) sl.push_back(0);
    DevBuf<uint32_t> ds, dn, dl;
    ds.reserve(map_size);
    dn.reserve(map_size);
    dl.reserve(sl.size());
    SGIPC_HIP(hipMemcpyAsync(ds.get(), ss.data(), sizeof(uint32_t) * map_size, hipMemcpyHostToDevice, s));
    SGIPC_HIP(hipMemcpyAsync(dn.get(), sn.data(), sizeof(uint32_t) * map_size, hipMemcpyHostToDevice, s));
    SGIPC_HIP(hipMemcpyAsync(dl.get(), sl.data(), sizeof(uint32_t) * sl.size(), hipMemcpyHostToDevice, s));
    sys->slot_of.reserve(n);
    SGIPC_HIP(hipMemcpyAsync(sys->slot_of.get(), r2p.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    sys->mas->m->init_neighbor(map_size, 0, total, ds.get(), dn.get(), dl.get());
    SGIPC_HIP(hipStreamSynchronize(s));
    sys->n_slots = map_size;
    sys->n_sys = map_size;
    sys->bs.reserve(3L * map_size);
    sys->xs.reserve(3L * map_size);
}

}  // namespace

extern "C" {

int sgipc_fem_tet_gradient_hessian(sgipc_ctx* ctx, const double* x, const uint32_t* tets, const double* dm_inv,
                                   const double* vol, const double* mu, const double* lam, int n_tets, double dt,
                                   int hess_offset, double* grad, double* tvals, int* trows, int* tcols)
{
    if (!ctx || n_tets < 0) return SGIPC_BAD_ARG;
    if (n_tets == 0) return SGIPC_EMPTY;
    if (!x || !tets || !dm_inv || !vol || !mu || !lam || !tvals || !trows || !tcols) return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        launch_tet_snh(ctx->c.stream, n_tets, x, tets, dm_inv, vol, mu, lam, dt, hess_offset, grad, tvals, trows,
                       tcols);
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int sgipc_barrier_pt_gradient_hessian(sgipc_ctx* ctx, const double* x, const int* pairs, const double* pair_x,
                                      int n_pairs, double dhat_sq, double kappa, int hess_offset, double* grad,
                                      double* tvals, int* trows, int* tcols)
{
    if (!ctx || n_pairs < 0) return SGIPC_BAD_ARG;
    if (n_pairs == 0) return SGIPC_EMPTY;
    if (!pairs || (!x && !pair_x) || !tvals || !trows || !tcols || !(dhat_sq > 0)) return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        launch_barrier_pt(ctx->c.stream, n_pairs, x, pairs, pair_x, dhat_sq, kappa, hess_offset, grad, tvals, trows,
                          tcols);
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int sgipc_fem_tri_gradient_hessian(sgipc_ctx* ctx, const double* x, const uint32_t* tris, const double* dm_inv2,
                                   const double* area, int n_tris, double k_stretch, double k_shear,
                                   double strain_rate, double dt, int hess_offset, double* grad, double* tvals,
                                   int* trows, int* tcols)
{
    if (!ctx || n_tris < 0) return SGIPC_BAD_ARG;
    if (n_tris == 0) return SGIPC_EMPTY;
    if (!x || !tris || !dm_inv2 || !area || !tvals || !trows || !tcols) return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        launch_tri_bw(ctx->c.stream, n_tris, x, tris, dm_inv2, area, k_stretch, k_shear, strain_rate, dt, hess_offset,
                      grad, tvals, trows, tcols);
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int sgipc_quad_bending_prepare(sgipc_ctx* ctx, const double* rest_x, const uint32_t* edges, const uint32_t* adj,
                               int n_edges, double* q16)
{
    if (!ctx || n_edges < 0) return SGIPC_BAD_ARG;
    if (n_edges == 0) return SGIPC_EMPTY;
    if (!rest_x || !edges || !adj || !q16) return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        launch_quad_bending_q(ctx->c.stream, n_edges, rest_x, edges, adj, q16);
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int sgipc_quad_bending_gradient_hessian(sgipc_ctx* ctx, const double* x, const double* rest_x, const uint32_t* edges,
                                        const uint32_t* adj, const double* q16, int n_edges, double k_bend, double dt,
                                        int hess_offset, double* grad, double* tvals, int* trows, int* tcols)
{
    if (!ctx || n_edges < 0) return SGIPC_BAD_ARG;
    if (n_edges == 0) return SGIPC_EMPTY;
    if (!x || !rest_x || !edges || !adj || !q16 || !tvals || !trows || !tcols) return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        launch_quad_bending(ctx->c.stream, n_edges, x, rest_x, edges, adj, q16, k_bend, dt, hess_offset, grad, tvals,
                            trows, tcols);
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int sgipc_abd_body_gradient_hessian(sgipc_ctx* ctx, int n_bodies, const double* q, const double* q_tilde,
                                    const double* mass10, const double* body_volume, double kappa, double dt,
                                    const int* body_fixed, double motor_speed, double motor_strength, double* grad,
                                    double* tvals, int* trows, int* tcols)
{
    if (!ctx || n_bodies < 0) return SGIPC_BAD_ARG;
    if (n_bodies == 0) return SGIPC_EMPTY;
    if (!q || !q_tilde || !mass10 || !body_volume || !tvals || !trows || !tcols) return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        launch_abd_body(ctx->c.stream, n_bodies, q, q_tilde, mass10, body_volume, kappa, dt, body_fixed, motor_speed,
                        motor_strength, grad, tvals, trows, tcols);
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int sgipc_contact_partition(sgipc_ctx* ctx, const double* vals, const int* rows, const int* cols, int n,
                            int n_abd_points, const int* point_body, const double* x_bar, int n_bodies,
                            const int* body_fixed, const int* fem_boundary_type, double* ovals, int* orows,
                            int* ocols, int capacity, int* out_n)
{
    if (!ctx || !out_n || n < 0 || n_abd_points < 0 || n_bodies < 0 || capacity < 0) return SGIPC_BAD_ARG;
    *out_n = 0;
    if (n == 0) return SGIPC_EMPTY;
    if (!vals || !rows || !cols || !ovals || !orows || !ocols) return SGIPC_BAD_ARG;
    if (n_abd_points > 0 && (!point_body || !x_bar)) return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        ContactMapper cm;
        const long total = cm.count(ctx->c.stream, rows, cols, n, n_abd_points);
        *out_n = (int)total;
        cm.write(ctx->c.stream, vals, rows, cols, n_abd_points, point_body, x_bar, n_bodies, body_fixed,
                 fem_boundary_type, ovals, orows, ocols, capacity);
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int sgipc_abd_jacobi_build(sgipc_ctx* ctx, const double* vals, const int* rows, const int* cols, int nnz,
                           int n_bodies, double* inv144)
{
    if (!ctx || nnz < 0 || n_bodies < 0) return SGIPC_BAD_ARG;
    if (n_bodies == 0) return SGIPC_EMPTY;
    if (!vals || !rows || !cols || !inv144) return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        DevBuf<int> bad;
        bad.reserve(1);
        SGIPC_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int), ctx->c.stream));
        launch_abd_jacobi_build(ctx->c.stream, vals, rows, cols, nnz, n_bodies, inv144, bad.get());
        int h = 0;
        SGIPC_HIP(hipMemcpyAsync(&h, bad.get(), sizeof(int), hipMemcpyDeviceToHost, ctx->c.stream));
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
        SGIPC_CHECK(h == 0, SGIPC_NOT_SPD, "singular affine-body block");
    });
}

int sgipc_abd_jacobi_apply(sgipc_ctx* ctx, const double* inv144, int n_bodies, const double* r, double* z)
{
    if (!ctx || n_bodies < 0) return SGIPC_BAD_ARG;
    if (n_bodies == 0) return SGIPC_EMPTY;
    if (!inv144 || !r || !z) return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        launch_abd_jacobi_apply<double>(ctx->c.stream, inv144, n_bodies, r, z, nullptr, nullptr);
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int sgipc_mass_diagonal(sgipc_ctx* ctx, const double* mass, int n, int hess_offset, double* tvals, int* trows,
                        int* tcols)
{
    if (!ctx || n < 0) return SGIPC_BAD_ARG;
    if (n == 0) return SGIPC_EMPTY;
    if (!mass || !tvals || !trows || !tcols) return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        launch_mass_diag(ctx->c.stream, n, mass, hess_offset, tvals, trows, tcols);
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int sgipc_convert_srbk(sgipc_ctx* ctx, const double* tvals, const int* trows, const int* tcols, int n_triplets,
                       double* ovals, int* orows, int* ocols, int capacity, int* out_nnz)
{
    if (!ctx || n_triplets < 0 || capacity < 0 || !out_nnz) return SGIPC_BAD_ARG;
    *out_nnz = 0;
    if (n_triplets == 0) return SGIPC_EMPTY;
    if (!tvals || !trows || !tcols || !ovals || !orows || !ocols) return SGIPC_BAD_ARG;
    const int st = run(ctx, [&] {
        Converter& cv = converter_of(ctx);
        int mx = cv.max_index(ctx->c.stream, trows, tcols, n_triplets);
        cv.last_unique = 0;
        *out_nnz = cv.convert(ctx->c.stream, tvals, trows, tcols, n_triplets, mx + 1, ovals, orows, ocols, capacity);
    });
    // a capacity failure reports the unique count the output needs
    if (st == SGIPC_CAPACITY && ctx->conv) *out_nnz = ctx->conv->last_unique;
    return st;
}

int sgipc_contact_detect(sgipc_ctx* ctx, const double* x, const double* rest_x, int n_verts, const uint32_t* faces,
                         int n_faces, const uint32_t* edges, int n_edges, const uint32_t* surf_verts, int n_surf,
                         const int* body_id, const int* boundary_type, double dhat_sq, int* pairs, int* mat_index,
                         int capacity, int* counts)
{
    if (!ctx || !counts || n_verts < 0 || n_faces < 0 || n_edges < 0 || n_surf < 0 || capacity < 0)
        return SGIPC_BAD_ARG;
    for (int i = 0; i < 5; ++i) counts[i] = 0;
    if (n_verts == 0 || (n_faces == 0 && n_edges < 2)) return SGIPC_EMPTY;
    if (!x || !(dhat_sq > 0) || (n_faces > 0 && (!faces || (n_surf > 0 && !surf_verts))) ||
        (n_edges > 0 && (!edges || !rest_x)) || (capacity > 0 && (!pairs || !mat_index)))
        return SGIPC_BAD_ARG;
    int total = 0;
    const int st = run(ctx, [&] {
        if (!ctx->contact) ctx->contact = std::make_unique<ContactDetector>();
        ContactCounts c = ctx->contact->detect(ctx->c.stream, x, rest_x, n_verts, faces, n_faces, edges, n_edges,
                                               surf_verts, n_surf, body_id, boundary_type, dhat_sq, pairs, mat_index,
                                               capacity);
        for (int i = 0; i < 5; ++i) counts[i] = c.c[i];
    });
    // a capacity failure reports the pair count the output needs
    if (st == SGIPC_CAPACITY && ctx->contact) total = ctx->contact->last_total;
    if (st == SGIPC_CAPACITY) counts[0] = total;
    return st;
}

int sgipc_barrier_gradient_hessian(sgipc_ctx* ctx, const double* x, const int* pairs, const int* mat_index,
                                   int n_pairs, const int* counts, double dhat_sq, double kappa, int hess_offset,
                                   double* grad, double* tvals, int* trows, int* tcols)
{
    if (!ctx || n_pairs < 0 || !counts) return SGIPC_BAD_ARG;
    if (n_pairs == 0) return SGIPC_EMPTY;
    if (!x || !pairs || !mat_index || !tvals || !trows || !tcols || !(dhat_sq > 0)) return SGIPC_BAD_ARG;
    if (counts[2] + counts[3] + counts[4] != n_pairs) return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        DevBuf<int> bad;
        bad.reserve(1);
        launch_fill<int>(ctx->c.stream, bad.get(), 0, 1);
        launch_barrier_pairs(ctx->c.stream, n_pairs, x, pairs, mat_index, counts[4], counts[3], dhat_sq, kappa,
                             hess_offset, grad, tvals, trows, tcols, bad.get());
        int h = 0;
        SGIPC_HIP(hipMemcpyAsync(&h, bad.get(), sizeof(int), hipMemcpyDeviceToHost, ctx->c.stream));
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
        SGIPC_CHECK(h == 0, SGIPC_BAD_ARG, "a pair in the smooth (mollified) encoding, which detection never writes");
    });
}

int sgipc_friction_prepare(sgipc_ctx* ctx, const double* x_t, const int* pairs, int n_pairs, double dhat_sq,
                           double kappa, double* lambda, double* dist_coord, double* tan_basis)
{
    if (!ctx || n_pairs < 0) return SGIPC_BAD_ARG;
    if (n_pairs == 0) return SGIPC_EMPTY;
    if (!x_t || !pairs || !lambda || !dist_coord || !tan_basis || !(dhat_sq > 0)) return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        DevBuf<int> bad;
        bad.reserve(1);
        launch_fill<int>(ctx->c.stream, bad.get(), 0, 1);
        launch_friction_prepare(ctx->c.stream, n_pairs, x_t, pairs, dhat_sq, kappa, lambda, dist_coord, tan_basis,
                                bad.get());
        int h = 0;
        SGIPC_HIP(hipMemcpyAsync(&h, bad.get(), sizeof(int), hipMemcpyDeviceToHost, ctx->c.stream));
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
        SGIPC_CHECK(h == 0, SGIPC_BAD_ARG, "a pair in the smooth (mollified) encoding, which detection never writes");
    });
}

int sgipc_friction_gradient_hessian(sgipc_ctx* ctx, const double* x, const double* x_t, const int* pairs,
                                    const int* mat_index, int n_pairs, const int* counts, const double* lambda,
                                    const double* dist_coord, const double* tan_basis, double eps_sq, double mu,
                                    int hess_offset, double* grad, double* tvals, int* trows, int* tcols)
{
    if (!ctx || n_pairs < 0 || !counts) return SGIPC_BAD_ARG;
    if (n_pairs == 0) return SGIPC_EMPTY;
    if (!x || !x_t || !pairs || !mat_index || !lambda || !dist_coord || !tan_basis || !tvals || !trows || !tcols ||
        !(eps_sq > 0))
        return SGIPC_BAD_ARG;
    if (counts[2] + counts[3] + counts[4] != n_pairs) return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        DevBuf<int> bad;
        bad.reserve(1);
        launch_fill<int>(ctx->c.stream, bad.get(), 0, 1);
        launch_friction(ctx->c.stream, n_pairs, x, x_t, pairs, mat_index, counts[4], counts[3], lambda, dist_coord,
                        tan_basis, eps_sq, mu, hess_offset, grad, tvals, trows, tcols, bad.get());
        int h = 0;
        SGIPC_HIP(hipMemcpyAsync(&h, bad.get(), sizeof(int), hipMemcpyDeviceToHost, ctx->c.stream));
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
        SGIPC_CHECK(h == 0, SGIPC_BAD_ARG, "a pair in the smooth (mollified) encoding, which detection never writes");
    });
}

int sgipc_ground_friction_prepare(sgipc_ctx* ctx, const double* x_t, const int* verts, int n, const double* normal,
                                  double offset, double dhat_sq, double kappa, double* lambda)
{
    if (!ctx || n < 0) return SGIPC_BAD_ARG;
    if (n == 0) return SGIPC_EMPTY;
    if (!x_t || !verts || !normal || !lambda || !(dhat_sq > 0)) return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        launch_ground_friction_prepare(ctx->c.stream, n, x_t, verts, normal, offset, dhat_sq, kappa, lambda);
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int sgipc_ground_friction_gradient_hessian(sgipc_ctx* ctx, const double* x, const double* x_t, const int* verts, int n,
                                           const double* normal, const double* lambda, double eps_sq, double mu,
                                           int hess_offset, double* grad, double* tvals, int* trows, int* tcols)
{
    if (!ctx || n < 0) return SGIPC_BAD_ARG;
    if (n == 0) return SGIPC_EMPTY;
    if (!x || !x_t || !verts || !normal || !lambda || !tvals || !trows || !tcols || !(eps_sq > 0)) return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        launch_ground_friction(ctx->c.stream, n, x, x_t, verts, normal, lambda, eps_sq, mu, hess_offset, grad, tvals,
                               trows, tcols);
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
}

int sgipc_ccd_self_step(sgipc_ctx* ctx, const double* x, const double* dir, double alpha, const uint32_t* faces,
                        int n_faces, const uint32_t* edges, int n_edges, const uint32_t* surf_verts, int n_surf,
                        const int* body_id, const int* boundary_type, double dhat_sq, double slackness, double* step,
                        int* n_candidates)
{
    if (!ctx || !step || n_faces < 0 || n_edges < 0 || n_surf < 0) return SGIPC_BAD_ARG;
    *step = 1.0;
    if (n_candidates) *n_candidates = 0;
    if (!x || !dir || !(dhat_sq > 0) || !(slackness > 0 && slackness < 1) || (n_faces > 0 && (!faces || !surf_verts)) ||
        (n_edges > 0 && !edges))
        return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        if (!ctx->contact) ctx->contact = std::make_unique<ContactDetector>();
        long nc = 0;
        *step = ctx->contact->ccd_step(ctx->c.stream, x, dir, alpha, faces, n_faces, edges, n_edges, surf_verts,
                                       n_surf, body_id, boundary_type, dhat_sq, slackness, &nc);
        if (n_candidates) *n_candidates = (int)nc;
    });
}

int sgipc_ground_step(sgipc_ctx* ctx, const double* x, const uint32_t* surf_verts, int n_surf, const double* normal,
                      double offset, const double* dir, double slackness, double* step)
{
    if (!ctx || !step || n_surf < 0) return SGIPC_BAD_ARG;
    *step = 1.0;
    if (n_surf == 0) return SGIPC_EMPTY;
    if (!x || !surf_verts || !normal || !dir || !(slackness > 0)) return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        StepReducer red;
        DevBuf<double> vals;
        *step = ground_step(ctx->c.stream, red, vals, x, surf_verts, n_surf, normal, offset, dir, slackness);
    });
}

int sgipc_injective_step(sgipc_ctx* ctx, const double* x, const uint32_t* tets, int n_tets, const double* dir,
                         double slackness, double error_rate, double* step)
{
    if (!ctx || !step || n_tets < 0) return SGIPC_BAD_ARG;
    *step = 1.0;
    if (n_tets == 0) return SGIPC_EMPTY;
    if (!x || !tets || !dir || !(slackness > 0 && slackness < 1)) return SGIPC_BAD_ARG;
    return run(ctx, [&] {
        StepReducer red;
        DevBuf<double> vals;
        *step = injective_step(ctx->c.stream, red, vals, x, tets, n_tets, dir, slackness, error_rate);
    });
}

int sgipc_mas_create(sgipc_ctx* ctx, sgipc_mas** out, int bank_size);
int sgipc_mas_destroy(sgipc_mas* mas);

int sgipc_system_create(sgipc_ctx* ctx, sgipc_system** out, const sgipc_scene_desc* desc, int bank_size)
{
    if (!ctx || !out || !desc) return SGIPC_BAD_ARG;
    *out = nullptr;
    const auto& d = *desc;
    if (d.n_verts <= 0) return SGIPC_EMPTY;
    if (d.n_tets < 0 || d.n_pairs < 0 || d.total_neighbor_num < 0) return SGIPC_BAD_ARG;
    if (!d.x || !d.mass || !d.nbr_start || !d.nbr_num || (d.total_neighbor_num > 0 && !d.nbr_list)) return SGIPC_BAD_ARG;
    if (d.n_tets > 0 && (!d.tets || !d.dm_inv || !d.vol || !d.mu || !d.lam)) return SGIPC_BAD_ARG;
    if (d.n_pairs > 0 && (!d.pairs || !(d.dhat_sq > 0))) return SGIPC_BAD_ARG;
    if (d.n_tris < 0 || d.n_hinges < 0) return SGIPC_BAD_ARG;
    if (d.n_tris > 0 && (!d.tris || !d.dm_inv2 || !d.area)) return SGIPC_BAD_ARG;
    if (d.n_hinges > 0 && (!d.rest_x || !d.hinge_edges || !d.hinge_adj)) return SGIPC_BAD_ARG;
    if (d.n_bodies < 0 || d.n_abd_points < 0) return SGIPC_BAD_ARG;
    if (d.n_bodies > 0 && (!d.q || !d.q_tilde || !d.mass10 || !d.body_volume)) return SGIPC_BAD_ARG;
    if (d.n_abd_points > 0 && (!d.point_body || !d.x_bar || d.n_bodies == 0)) return SGIPC_BAD_ARG;
    if (d.n_bodies > 0 && d.n_pairs > 0 && !d.pair_x) return SGIPC_BAD_ARG;
    if (bank_size != 16 && bank_size != 32) return SGIPC_BAD_ARG;
    if (d.part_id && d.n_bodies > 0) return SGIPC_BAD_ARG;  // partitions cover FEM-only systems
    if (d.detect_contacts) {
        if (d.n_bodies > 0 || d.n_faces < 0 || d.n_edges < 0 || d.n_surf < 0 || !(d.dhat_sq > 0)) return SGIPC_BAD_ARG;
        if ((d.n_faces > 0 && (!d.faces || !d.surf_verts)) || (d.n_edges > 0 && (!d.edges || !d.rest_x)))
            return SGIPC_BAD_ARG;
    }
    auto* sys = new (std::nothrow) sgipc_system();
    if (!sys) return SGIPC_HIP_ERROR;
    sys->ctx = ctx;
    sys->d = d;
    sys->B = bank_size;
    int st = sgipc_mas_create(ctx, &sys->mas, bank_size);
    if (st == SGIPC_OK)
        st = run(ctx, [&] {
            sys->n_sys = 4 * d.n_bodies + d.n_verts;
            if (d.part_id) {
                init_slots(sys);
            } else {
                sys->mas->m->init_neighbor(d.n_verts, d.n_abd_points, d.total_neighbor_num, d.nbr_start, d.nbr_num,
                                           d.nbr_list);
            }
            sys->mas->m->init_matrix();
            sys->n_trip = 10L * d.n_bodies + 10L * d.n_tets + 10L * d.n_hinges + 6L * d.n_tris + d.n_verts;
            if (d.n_hinges > 0) {
                sys->hinge_q.reserve(16L * d.n_hinges);
                launch_quad_bending_q(ctx->c.stream, d.n_hinges, d.rest_x, d.hinge_edges, d.hinge_adj,
                                      sys->hinge_q.get());
            }
            long contact = d.detect_contacts ? 0 : 16L * d.n_pairs;
            if (d.detect_contacts) {
                sys->dcap = d.contact_capacity > 0 ? d.contact_capacity : 8 * (d.n_surf + d.n_edges) + 64;
                sys->dpairs.reserve(4L * sys->dcap);
                sys->dmat.reserve(sys->dcap);
                sys->dbad.reserve(1);
            }
            if (d.n_bodies > 0) {
                sys->cvals.reserve((size_t)(contact > 0 ? contact : 1) * 9);
                sys->crows.reserve(contact > 0 ? contact : 1);
                sys->ccols.reserve(contact > 0 ? contact : 1);
                contact *= 4;  // first guess of the mapped blocks; grown on demand
            }
            sys->arena_cap = sys->n_trip + contact;
            SGIPC_CHECK(sys->arena_cap < (1L << 31), SGIPC_CAPACITY, "triplet count exceeds 2^31");
            sys->tvals.reserve((size_t)sys->arena_cap * 9);
            sys->trows.reserve(sys->arena_cap);
            sys->tcols.reserve(sys->arena_cap);
            sys->capacity = (int)sys->arena_cap;
            sys->bvals.reserve((size_t)sys->capacity * 9);
            sys->brows.reserve(sys->capacity);
            sys->bcols.reserve(sys->capacity);
            sys->ws = new SolveWorkspace();
            for (auto& e : sys->ev) SGIPC_HIP(hipEventCreate(&e));
        });
    if (st != SGIPC_OK) {
        if (sys->mas) sgipc_mas_destroy(sys->mas);
        delete sys->ws;
        delete sys;
        return st;
    }
    *out = sys;
    return SGIPC_OK;
}

int sgipc_system_destroy(sgipc_system* sys)
{
    if (!sys) return SGIPC_BAD_ARG;
    int st = run(sys->ctx, [&] {
        for (auto& e : sys->ev)
            if (e) SGIPC_HIP(hipEventDestroy(e));
        delete sys->ws;
        sys->ws = nullptr;
        sys->tvals.release();
        sys->trows.release();
        sys->tcols.release();
        sys->bvals.release();
        sys->brows.release();
        sys->bcols.release();
        sys->hinge_q.release();
    });
    if (sys->mas) sgipc_mas_destroy(sys->mas);
    delete sys;
    return st;
}

}  // extern "C"

namespace {

// Whether this solve uses the learned correction.  Automatic mode measures
// both variants once, then keeps the one with fewer iterations and measures
// the other again later (the matrix drifts over Newton steps): after 8
// solves, with the gap doubling every time the learned level loses again and
// reset to 8 when it wins.
bool use_learned(const sgipc_system* sys)
{
    if (!sys->net || sys->learned_mode == 0) return false;
    if (sys->learned_mode == 1) return true;
    if (sys->plain_it < 0) return false;
    if (sys->learned_it < 0) return true;
    const bool better = sys->learned_it < sys->plain_it;
    return sys->n_solves == sys->next_probe ? !better : better;
}

int system_solve(sgipc_system* sys, const double* b, double* x, double* grad, const sgipc_pcg_opts* opts,
                 sgipc_pcg_stats* stats, sgipc_solve_timing* timing, const PcgOptions* record)
{
    if (!sys || !b || !x) return SGIPC_BAD_ARG;
    sgipc_pcg_opts o;
    sgipc_pcg_default_opts(&o);
    if (opts) o = *opts;
    if (o.precision < SGIPC_FP64 || o.precision > SGIPC_FP32_IR) return SGIPC_BAD_ARG;
    const bool slots = sys->n_slots > 0;
    // the iteration cap stays 0.3 x the vertex-space unknowns in slot space
    if (slots && o.max_iter <= 0) o.max_iter = std::max(2, (int)(0.3 * 3.0 * (double)sys->d.n_verts));
    int status = SGIPC_OK;
    int st = run(sys->ctx, [&] {
        const auto& d = sys->d;
        hipStream_t s = sys->ctx->c.stream;
        auto t0 = std::chrono::steady_clock::now();
        SGIPC_HIP(hipEventRecord(sys->ev[0], s));
        // ---- assembly (arena order: contacts, bodies, tets, hinges, triangles, mass)
        const int nb = d.n_bodies;
        const int off0 = 4 * nb;  // FEM block rows start after the bodies
        if (grad) launch_fill<double>(s, grad, 0.0, 3L * sys->n_sys);
        long off = 0;
        const int* mas_pairs = d.pairs;
        int mas_npairs = d.n_pairs;
        if (d.detect_contacts) {
            // the pairs of the current positions (grown once when short), then
            // every pair type's barrier into the contact arena
            ContactCounts cc;
            for (int attempt = 0;; ++attempt) {
                try {
                    cc = sys->det.detect(s, d.x, d.rest_x, d.n_verts, d.faces, d.n_faces, d.edges, d.n_edges,
                                         d.surf_verts, d.n_surf, d.body_id, d.boundary_type, d.dhat_sq,
                                         sys->dpairs.get(), sys->dmat.get(), sys->dcap);
                    break;
                } catch (const Error& e) {
                    if (e.code != SGIPC_CAPACITY || attempt > 0 || sys->det.last_total <= sys->dcap) throw;
                    sys->dcap = sys->det.last_total + sys->det.last_total / 4 + 64;
                    sys->dpairs.reserve(4L * sys->dcap);
                    sys->dmat.reserve(sys->dcap);
                }
            }
            const long ct = 16L * cc.c[4] + 9L * cc.c[3] + 4L * cc.c[2];
            if (ct + sys->n_trip > sys->arena_cap) {
                sys->arena_cap = ct + sys->n_trip;
                SGIPC_CHECK(sys->arena_cap < (1L << 31), SGIPC_CAPACITY, "triplet count exceeds 2^31");
                sys->tvals.reserve((size_t)sys->arena_cap * 9);
                sys->trows.reserve(sys->arena_cap);
                sys->tcols.reserve(sys->arena_cap);
                sys->capacity = (int)sys->arena_cap;
                sys->bvals.reserve((size_t)sys->capacity * 9);
                sys->brows.reserve(sys->capacity);
                sys->bcols.reserve(sys->capacity);
            }
            launch_fill<int>(s, sys->dbad.get(), 0, 1);
            launch_barrier_pairs(s, cc.c[0], d.x, sys->dpairs.get(), sys->dmat.get(), cc.c[4], cc.c[3], d.dhat_sq,
                                 d.kappa, 0, grad, sys->tvals.get(), sys->trows.get(), sys->tcols.get(),
                                 sys->dbad.get());
            off += ct;
            mas_pairs = sys->dpairs.get();
            mas_npairs = cc.c[0];
            sys->n_detected = cc.c[0];
        } else if (nb == 0) {
            launch_barrier_pt(s, d.n_pairs, d.x, d.pairs, d.pair_x, d.dhat_sq, d.kappa, 0, grad, sys->tvals.get(),
                              sys->trows.get(), sys->tcols.get());
            off += 16L * d.n_pairs;
        } else {
            // point-level contact blocks and gradient, then the class mapping
            // into the arena and J^T g into the body rows
            double* gp = nullptr;
            const long n_points = (long)d.n_abd_points + d.n_verts;
            if (grad && d.n_pairs > 0) {
                sys->gpt.reserve(3 * n_points);
                launch_fill<double>(s, sys->gpt.get(), 0.0, 3 * n_points);
                gp = sys->gpt.get();
            }
            launch_barrier_pt(s, d.n_pairs, d.x, d.pairs, d.pair_x, d.dhat_sq, d.kappa, 0, gp, sys->cvals.get(),
                              sys->crows.get(), sys->ccols.get());
            const long m = sys->cmap.count(s, sys->crows.get(), sys->ccols.get(), 16L * d.n_pairs, d.n_abd_points);
            if (m + sys->n_trip > sys->arena_cap) {
                sys->arena_cap = m + sys->n_trip;
                SGIPC_CHECK(sys->arena_cap < (1L << 31), SGIPC_CAPACITY, "triplet count exceeds 2^31");
                sys->tvals.reserve((size_t)sys->arena_cap * 9);
                sys->trows.reserve(sys->arena_cap);
                sys->tcols.reserve(sys->arena_cap);
                sys->capacity = (int)sys->arena_cap;
                sys->bvals.reserve((size_t)sys->capacity * 9);
                sys->brows.reserve(sys->capacity);
                sys->bcols.reserve(sys->capacity);
            }
            sys->cmap.write(s, sys->cvals.get(), sys->crows.get(), sys->ccols.get(), d.n_abd_points, d.point_body,
                            d.x_bar, nb, d.body_fixed, d.boundary_type, sys->tvals.get(), sys->trows.get(),
                            sys->tcols.get(), m);
            off += m;
            launch_abd_body(s, nb, d.q, d.q_tilde, d.mass10, d.body_volume, d.kappa_abd, d.dt, d.body_fixed,
                            d.motor_speed, d.motor_strength, grad, sys->tvals.get() + 9 * off, sys->trows.get() + off,
                            sys->tcols.get() + off);
            if (gp)
                launch_point_gradient(s, (int)n_points, d.n_abd_points, gp, d.point_body, d.x_bar, d.body_fixed, nb,
                                      grad);
            off += 10L * nb;
        }
        double* gf = grad ? grad + 3L * off0 : nullptr;
        launch_tet_snh(s, d.n_tets, d.x, d.tets, d.dm_inv, d.vol, d.mu, d.lam, d.dt, off0, gf,
                       sys->tvals.get() + 9 * off, sys->trows.get() + off, sys->tcols.get() + off);
        off += 10L * d.n_tets;
        launch_quad_bending(s, d.n_hinges, d.x, d.rest_x, d.hinge_edges, d.hinge_adj, sys->hinge_q.get(), d.k_bend,
                            d.dt, off0, gf, sys->tvals.get() + 9 * off, sys->trows.get() + off, sys->tcols.get() + off);
        off += 10L * d.n_hinges;
        launch_tri_bw(s, d.n_tris, d.x, d.tris, d.dm_inv2, d.area, d.k_stretch, d.k_shear, d.strain_rate, d.dt, off0,
                      gf, sys->tvals.get() + 9 * off, sys->trows.get() + off, sys->tcols.get() + off);
        off += 6L * d.n_tris;
        // Dirichlet vertices: every block before the mass diagonal with a
        // fixed FEM end is zeroed, and their gradient rows
        launch_dirichlet_zero(s, off, sys->trows.get(), sys->tcols.get(), off0, d.boundary_type, sys->tvals.get());
        if (gf) launch_dirichlet_zero_rows(s, d.n_verts, d.boundary_type, gf);
        launch_mass_diag(s, d.n_verts, d.mass, off0, sys->tvals.get() + 9 * off, sys->trows.get() + off,
                         sys->tcols.get() + off);
        off += d.n_verts;
        const double* bp = b;
        double* xp = x;
        if (slots) {
            // vertex rows and columns renamed to slots; contact pairs for the
            // MAS connectivity likewise; b placed at the slots
            launch_slot_triplets(s, off, d.n_verts, sys->slot_of.get(), sys->tvals.get(), sys->trows.get(),
                                 sys->tcols.get());
            if (mas_npairs > 0) {
                sys->spairs.reserve(4L * mas_npairs);
                launch_slot_pairs(s, mas_npairs, sys->slot_of.get(), mas_pairs, sys->spairs.get());
                mas_pairs = sys->spairs.get();
            }
            launch_slot_scatter(s, d.n_verts, sys->n_slots, sys->slot_of.get(), b, sys->bs.get());
            bp = sys->bs.get();
            xp = sys->xs.get();
        }
        SGIPC_HIP(hipEventRecord(sys->ev[1], s));
        // ---- triplets -> unique upper BCOO; meanwhile a helper thread builds
        // the MAS hierarchy (connectivity and contact pairs only) on a second
        // stream that starts after the pairs are written
        SolveWorkspace& ws = *sys->ws;
        ws.layout_side.begin(s);
        std::exception_ptr hier_err;
        std::thread hier([&] {
            try {
                sys->ctx->c.activate();
                sys->mas->m->prebuild_hierarchy(ws.layout_side.s, mas_pairs, mas_npairs);
            } catch (...) {
                hier_err = std::current_exception();
            }
        });
        try {
            sys->nnz = sys->conv.convert(s, sys->tvals.get(), sys->trows.get(), sys->tcols.get(), off, sys->n_sys,
                                         sys->bvals.get(), sys->brows.get(), sys->bcols.get(), sys->capacity);
        } catch (...) {
            hier.join();
            throw;
        }
        hier.join();
        if (hier_err) std::rethrow_exception(hier_err);
        SGIPC_HIP(hipEventRecord(sys->ev[2], s));
        // ---- preconditioners: MAS on the FEM rows (current contacts), block-Jacobi on the bodies
        // mixed precision failed on an earlier solve: FP64 from the start
        // until the next probe (not while collecting training samples)
        const bool skip32 = o.precision != SGIPC_FP64 && record == nullptr && sys->fp32_failed &&
                            sys->n_solves < sys->fp32_probe;
        if (skip32) o.precision = SGIPC_FP64;
        const int mprec = (o.precision == SGIPC_FP64) ? SGIPC_FP64 : SGIPC_FP32_MAS;
        const bool learned = record == nullptr && use_learned(sys);
        sys->mas->m->keep_graph = sys->net != nullptr || record != nullptr;
        // the PCG matrix serves as coarse mode 2's fine operator when the MAS
        // covers every row; behind body rows the MAS builds its own
        sys->mas->m->build_own_fine = off0 > 0;
        // the PCG's SpMV layout is built by a helper thread on a second stream
        // while this thread builds the preconditioners (both only read the
        // BCOO; each build waits on its own stream's read-backs)
        ws.prebuilt = false;
        ws.layout_side.begin(s);
        std::exception_ptr layout_err;
        std::thread layout([&] {
            try {
                sys->ctx->c.activate();
                ws.full.build(ws.layout_side.s, sys->bvals.get(), sys->brows.get(), sys->bcols.get(), sys->nnz,
                              sys->n_sys, o.precision != SGIPC_FP64);
            } catch (...) {
                layout_err = std::current_exception();
            }
        });
        try {
            sys->mas->m->set_bcoo(sys->bvals.get(), sys->brows.get(), sys->bcols.get(), sys->nnz, off0, mas_pairs,
                                  mas_npairs, mprec);
            if (nb > 0) sys->abdj.build(s, sys->bvals.get(), sys->brows.get(), sys->bcols.get(), sys->nnz, nb);
        } catch (...) {
            layout.join();
            throw;
        }
        layout.join();
        if (layout_err) std::rethrow_exception(layout_err);
        ws.layout_side.end(s);
        ws.prebuilt = true;
        struct ClearPrebuilt {
            SolveWorkspace& w;
            ~ClearPrebuilt() { w.prebuilt = false; }
        } clear_prebuilt{ws};
        SGIPC_HIP(hipEventRecord(sys->ev[3], s));
        // ---- PCG
        sgipc_pcg_stats pst{};
        SubsystemPrec prec;
        prec.mas = sys->mas->m.get();
        prec.mas_row_offset = off0;
        prec.abd = nb > 0 ? &sys->abdj : nullptr;
        PcgOptions extra;
        if (record) extra = *record;
        extra.flexible = learned;
        sys->mas->m->net_ = learned ? mpnn_of(sys->net) : nullptr;
        if (learned && !sys->mas->m->learned_active()) sys->mas->m->net_ = nullptr;
        bool used = sys->mas->m->net_ != nullptr;
        extra.flexible = used;
        // automatic mode: a run with the learned correction gets at most twice
        // the plain variant's iterations; one that does not converge there has
        // lost, and the call solves again without it
        const bool capped = used && sys->learned_mode == 2 && sys->plain_it > 0;
        sgipc_pcg_opts oc = o;
        if (capped) {
            const int cap = 2 * sys->plain_it + 16;
            if (oc.max_iter <= 0 || oc.max_iter > cap) oc.max_iter = cap;
        }
        auto run_pcg = [&](const sgipc_pcg_opts& oo) {
            try {
                return pcg_solve_impl(sys->ctx->c, *sys->ws, sys->bvals.get(), sys->brows.get(), sys->bcols.get(),
                                      sys->nnz, sys->n_sys, prec, bp, xp, oo, &pst, &extra);
            } catch (...) {
                sys->mas->m->net_ = nullptr;
                throw;
            }
        };
        status = run_pcg(capped ? oc : o);
        bool abandoned = false;
        int spent = 0;  // iterations of runs that failed and were redone
        // a learned run that missed its cap, or whose correction made the
        // preconditioner indefinite (r.z <= 0, in either mode), has lost:
        // the same call solves again without it
        if (used && ((capped && status != SGIPC_OK) || status == SGIPC_NOT_SPD)) {
            spent += pst.iterations;
            sys->learned_it = std::numeric_limits<int>::max() / 2;
            sys->mas->m->net_ = nullptr;
            extra.flexible = false;
            used = false;
            abandoned = true;
            status = run_pcg(o);
        }
        sys->mas->m->net_ = nullptr;
        // coarse mode 2 is positive definite only for smoothing weights below
        // 2 / lambda_max(M0 A): a run of the two-grid cycle that meets
        // r.z <= 0 is redone with half the weight (at most twice, never below
        // kOmegaFloor; a weight that cures it is kept for later solves), and
        // then with the level-1 cycle alone, the weight restored (halving did
        // not cure it, so the weight was not the cause)
        Mas* mm = sys->mas->m.get();
        constexpr double kOmegaFloor = 0.05;
        auto recover_spd = [&](const sgipc_pcg_opts& oo) {
            if (!(status == SGIPC_NOT_SPD && mm->last_path() == 2 && !used)) return;
            const double w0 = mm->omega0;
            for (int tries = 0; tries < 2 && status == SGIPC_NOT_SPD && mm->omega0 * 0.5 >= kOmegaFloor; ++tries) {
                spent += pst.iterations;
                mm->omega0 *= 0.5;
                status = run_pcg(oo);
            }
            if (status == SGIPC_NOT_SPD) {
                spent += pst.iterations;
                mm->omega0 = w0;
                mm->coarse_mode = 1;
                try {
                    status = run_pcg(oo);
                } catch (...) {
                    mm->coarse_mode = 2;
                    throw;
                }
                mm->coarse_mode = 2;
            }
        };
        recover_spd(o);
        // a mixed-precision solve that still failed because the FP32
        // operators do not resolve the system (an indefinite FP32/FP16
        // preconditioner, or refinement passes that stopped reducing the FP64
        // residual before the iteration cap) is redone in FP64 with the
        // preconditioner rebuilt in FP64, with the same recovery
        int fallback = 0;
        const int cap_it = o.max_iter > 0 ? o.max_iter : (int)(0.3 * 3.0 * (double)sys->n_sys);
        const bool precision_failed =
            status == SGIPC_NOT_SPD || (status == SGIPC_MAXITER && pst.iterations < cap_it - 1);
        if (precision_failed && mprec != SGIPC_FP64 && record == nullptr) {
            sgipc_pcg_opts o64 = o;
            o64.precision = SGIPC_FP64;
            mm->net_ = nullptr;
            extra.flexible = false;
            used = false;
            mm->set_bcoo(sys->bvals.get(), sys->brows.get(), sys->bcols.get(), sys->nnz, off0, mas_pairs, mas_npairs,
                         SGIPC_FP64);
            spent += pst.iterations;
            status = run_pcg(o64);
            recover_spd(o64);
            fallback = 1;
            if (sys->fp32_failed) sys->fp32_gap *= 2;
            sys->fp32_failed = true;
            sys->fp32_probe = sys->n_solves + sys->fp32_gap;
        } else if (mprec != SGIPC_FP64 && record == nullptr && status == SGIPC_OK) {
            sys->fp32_failed = false;  // mixed precision works (again)
            sys->fp32_gap = 8;
        }
        if (skip32) fallback = 2;
        pst.iterations += spent;
        if (slots) launch_slot_gather(s, d.n_verts, sys->slot_of.get(), sys->xs.get(), x);
        pst.fp64_fallback = fallback;
        pst.learned = used ? 1 : (abandoned ? 2 : 0);
        pst.coarse_mode = mm->last_path();
        pst.level2_exact = mm->level2_exact() ? 1 : 0;
        pst.omega = mm->omega0;
        if (record == nullptr && status == SGIPC_OK) {
            (used ? sys->learned_it : sys->plain_it) = pst.iterations;
            if (sys->learned_mode == 2 && sys->plain_it >= 0 && sys->learned_it >= 0 &&
                sys->n_solves >= sys->next_probe) {
                sys->probe_gap = sys->learned_it < sys->plain_it ? 8 : 2 * sys->probe_gap;
                sys->next_probe = sys->n_solves + sys->probe_gap;
            }
            ++sys->n_solves;
        }
        SGIPC_HIP(hipEventRecord(sys->ev[4], s));
        SGIPC_HIP(hipEventSynchronize(sys->ev[4]));
        auto t1 = std::chrono::steady_clock::now();
        if (stats) {
            *stats = pst;
            stats->ms_total = std::chrono::duration<double, std::milli>(t1 - t0).count();
        }
        if (timing) {
            float ms[4];
            for (int i = 0; i < 4; ++i) SGIPC_HIP(hipEventElapsedTime(&ms[i], sys->ev[i], sys->ev[i + 1]));
            timing->ms_assemble = ms[0];
            timing->ms_convert = ms[1];
            timing->ms_mas_build = ms[2];
            timing->ms_pcg = ms[3];
            timing->ms_total = std::chrono::duration<double, std::milli>(t1 - t0).count();
            timing->nnz = sys->nnz;
            timing->n_triplets = (int)off;
        }
    });
    return st != SGIPC_OK ? st : status;
}

}  // namespace

// Objective of the Newton step at positions x (GIPC::computeEnergy,
// StiffGIPC/GIPC.cu:7373-7870 restricted to the terms this system assembles):
// 1/2 |x - x~|^2_M (when x_tilde is given) + the elastic terms + the barrier
// of the pairs at x (detected there when the scene detects its contacts).
double system_energy(sgipc_system* sys, const double* x, const double* x_tilde)
{
    const auto& d = sys->d;
    hipStream_t s = sys->ctx->c.stream;
    SGIPC_CHECK(d.n_bodies == 0, SGIPC_BAD_ARG, "the energy of affine bodies is not evaluated");
    EnergySum& e = sys->esum;
    double E = e.inertia(s, d.n_verts, x, x_tilde, d.mass);
    E += e.tets(s, d.n_tets, x, d.tets, d.dm_inv, d.vol, d.mu, d.lam, d.dt);
    E += e.tris(s, d.n_tris, x, d.tris, d.dm_inv2, d.area, d.k_stretch, d.k_shear, d.strain_rate, d.dt);
    E += e.bending(s, d.n_hinges, x, d.rest_x, d.hinge_edges, d.hinge_adj, sys->hinge_q.get(), d.k_bend, d.dt);
    if (d.detect_contacts) {
        if (sys->lcap == 0) {
            sys->lcap = sys->dcap;
            sys->lpairs.reserve(4L * sys->lcap);
            sys->lmat.reserve(sys->lcap);
        }
        ContactCounts cc;
        for (int attempt = 0;; ++attempt) {
            try {
                cc = sys->det.detect(s, x, d.rest_x, d.n_verts, d.faces, d.n_faces, d.edges, d.n_edges, d.surf_verts,
                                     d.n_surf, d.body_id, d.boundary_type, d.dhat_sq, sys->lpairs.get(),
                                     sys->lmat.get(), sys->lcap);
                break;
            } catch (const Error& err) {
                if (err.code != SGIPC_CAPACITY || attempt > 0 || sys->det.last_total <= sys->lcap) throw;
                sys->lcap = sys->det.last_total + sys->det.last_total / 4 + 64;
                sys->lpairs.reserve(4L * sys->lcap);
                sys->lmat.reserve(sys->lcap);
            }
        }
        E += e.barrier(s, cc.c[0], x, sys->lpairs.get(), nullptr, d.dhat_sq, d.kappa);
    } else {
        E += e.barrier(s, d.n_pairs, x, d.pairs, d.pair_x, d.dhat_sq, d.kappa);
    }
    return E;
}

extern "C" {

int sgipc_system_solve(sgipc_system* sys, const double* b, double* x, double* grad, const sgipc_pcg_opts* opts,
                       sgipc_pcg_stats* stats, sgipc_solve_timing* timing)
{
    return system_solve(sys, b, x, grad, opts, stats, timing, nullptr);
}

int sgipc_system_set_learned(sgipc_system* sys, sgipc_mpnn* net, int mode)
{
    if (!sys || mode < 0 || mode > 2 || (mode != 0 && !net)) return SGIPC_BAD_ARG;
    sys->net = mode ? net : nullptr;
    sys->learned_mode = mode;
    sys->plain_it = sys->learned_it = -1;
    sys->n_solves = 0;
    sys->next_probe = sys->probe_gap = 8;
    return SGIPC_OK;
}

int sgipc_system_collect(sgipc_system* sys, const double* b, const sgipc_pcg_opts* opts, int stride,
                         int max_samples, sgipc_mpnn_data** out)
{
    if (!sys || !b || !out || stride < 1 || max_samples < 1) return SGIPC_BAD_ARG;
    *out = nullptr;
    const long n3 = 3L * sys->n_sys;
    DevBuf<double> rec, xtmp, ztmp;
    int count = 0;
    int st = run(sys->ctx, [&] {
        rec.reserve((size_t)max_samples * n3);
        xtmp.reserve(n3);
        ztmp.reserve(n3);
    });
    if (st != SGIPC_OK) return st;
    PcgOptions r;
    r.record = rec.get();
    r.record_stride = stride;
    r.record_cap = max_samples;
    r.record_count = &count;
    // residuals are recorded by the single-run solver: FP32_IR collects with
    // the FP32 MAS apply and the FP64 SpMV
    sgipc_pcg_opts o;
    sgipc_pcg_default_opts(&o);
    if (opts) o = *opts;
    if (o.precision == SGIPC_FP32_IR) o.precision = SGIPC_FP32_MAS;
    st = system_solve(sys, b, xtmp.get(), nullptr, &o, nullptr, nullptr, &r);
    if (st != SGIPC_OK && st != SGIPC_MAXITER) return st;
    sgipc_mpnn_data* d = mpnn_data_new(sys->ctx);
    st = run(sys->ctx, [&] {
        hipStream_t s = sys->ctx->c.stream;
        Mas& m = *sys->mas->m;
        SGIPC_CHECK(m.level1_graph().n > 0, SGIPC_EMPTY, "the MAS has no level-1 graph (one level)");
        SGIPC_CHECK(sys->d.n_bodies == 0, SGIPC_BAD_ARG, "training samples need a FEM-only system");
        mpnn_data_graph(d).copy_from(s, m.level1_graph());
        const long rows = 3L * m.level1_graph().n;
        DevBuf<float> X, q;
        X.reserve((size_t)rows * 3);
        q.reserve(rows);
        for (int k = 0; k < count; ++k) {
            m.learned_sample(s, rec.get() + (size_t)k * n3, ztmp.get(), X.get(), q.get());
            mpnn_data_append(d, X.get(), q.get(), 1);
        }
        SGIPC_HIP(hipStreamSynchronize(s));
    });
    if (st != SGIPC_OK) {
        (void)sgipc_mpnn_data_destroy(d);
        return st;
    }
    *out = d;
    return SGIPC_OK;
}

int sgipc_system_bcoo(sgipc_system* sys, const double** vals, const int** rows, const int** cols, int* nnz)
{
    if (!sys || !vals || !rows || !cols || !nnz) return SGIPC_BAD_ARG;
    *vals = sys->bvals.get();
    *rows = sys->brows.get();
    *cols = sys->bcols.get();
    *nnz = sys->nnz;
    return SGIPC_OK;
}

sgipc_mas* sgipc_system_mas(sgipc_system* sys) { return sys ? sys->mas : nullptr; }

int sgipc_system_energy(sgipc_system* sys, const double* x, const double* x_tilde, double* energy)
{
    if (!sys || !x || !energy) return SGIPC_BAD_ARG;
    return run(sys->ctx, [&] { *energy = system_energy(sys, x, x_tilde); });
}

int sgipc_system_line_search(sgipc_system* sys, const double* dir, const double* x_tilde, double alpha,
                             double slackness, double* x_out, double* alpha_out, double* energies, int* stopped)
{
    if (!sys || !dir || !x_out || !alpha_out || !(alpha > 0) || !(slackness > 0 && slackness < 1)) return SGIPC_BAD_ARG;
    return run(sys->ctx, [&] {
        const auto& d = sys->d;
        hipStream_t s = sys->ctx->c.stream;
        const long n3 = 3L * d.n_verts;
        double a = alpha;
        // feasible step: no contact along the path, no inverted tet
        if (d.detect_contacts) {
            if (!sys->ctx->contact) sys->ctx->contact = std::make_unique<ContactDetector>();
            a = std::min(a, sys->ctx->contact->ccd_step(s, d.x, dir, a, d.faces, d.n_faces, d.edges, d.n_edges,
                                                        d.surf_verts, d.n_surf, d.body_id, d.boundary_type,
                                                        d.dhat_sq, slackness, nullptr));
        }
        if (d.n_tets > 0)
            a = std::min(a, injective_step(s, sys->stepred, sys->stepvals, d.x, d.tets, d.n_tets, dir, slackness,
                                           1e-20));
        const double E0 = system_energy(sys, d.x, x_tilde);
        auto trial = [&](double t) {
            SGIPC_HIP(hipMemcpyAsync(x_out, d.x, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
            launch_axpy_f64(s, -t, dir, x_out, n3);
            return system_energy(sys, x_out, x_tilde);
        };
        double E = trial(a);
        int k = 0;
        // Armijo with c1 = 0 and at most 8 halvings (GIPC.cu:10828-10920)
        while (E > E0 && k <= 8) {
            a *= 0.5;
            ++k;
            E = trial(a);
        }
        int st = 0;
        if (k > 8) {
            SGIPC_HIP(hipMemcpyAsync(x_out, d.x, sizeof(double) * n3, hipMemcpyDeviceToDevice, s));
            SGIPC_HIP(hipStreamSynchronize(s));
            st = 1;
        }
        *alpha_out = a;
        if (energies) {
            energies[0] = E0;
            energies[1] = E;
        }
        if (stopped) *stopped = st;
    });
}

int sgipc_system_contacts(sgipc_system* sys, const int** pairs, const int** mat_index, int* n_pairs)
{
    if (!sys || !pairs || !mat_index || !n_pairs) return SGIPC_BAD_ARG;
    *pairs = sys->d.detect_contacts ? sys->dpairs.get() : sys->d.pairs;
    *mat_index = sys->d.detect_contacts ? sys->dmat.get() : nullptr;
    *n_pairs = sys->d.detect_contacts ? sys->n_detected : sys->d.n_pairs;
    return SGIPC_OK;
}

}  // extern "C"
