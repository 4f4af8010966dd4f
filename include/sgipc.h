/*
 * sgipc.h -- C ABI of the MI355X (gfx950) Newton linear-solve library.
 *
 * Drop-in boundary for StiffGIPC's linear-solve hot path.  Every entry point
 * takes plain pointers and sizes; array arguments are DEVICE pointers unless
 * the comment says "host".  Conventions shared with the reference:
 *   - a 3x3 block is 9 doubles (or floats), column-major like Eigen::Matrix3d:
 *     v[c*3 + r] = M(r, c);
 *   - vectors are contiguous doubles of length 3*rows (double3 aliasing);
 *   - the assembled matrix is upper-triangular symmetric BCOO (row <= col),
 *     sorted by (row, col), unique (the output of Converter::srbk_convert);
 *   - contact pairs are int4 in the reference encoding, point-triangle =
 *     (-v-1, t0, t1, t2).
 *
 * Error behaviour: the reference aborts on any CUDA error (CUDA_SAFE_CALL,
 * StiffGIPC/cuda_tools/cuda_tools.h:36-46).  Here every function returns an
 * int status and never aborts; sgipc_last_error() gives the message.
 * Threading: one host thread per context; all work is ordered on the
 * context's HIP stream and every function returns after its results are
 * complete (synchronous semantics, like the reference's callers expect).
 * Inside a call the library may run independent build work on streams of
 * its own (default priority), ordered against the context's stream by
 * events, and a system solve builds its SpMV layout on a helper thread that
 * is joined before the call goes on; nothing outlives the call.
 */
#ifndef SGIPC_H
#define SGIPC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SGIPC_ABI_VERSION 2

enum sgipc_status {
    SGIPC_OK = 0,
    SGIPC_EMPTY = 1,      /* nothing to do (n = 0 or nnz = 0); outputs untouched  */
    SGIPC_NOT_SPD = 2,    /* p.Ap <= 0 met during PCG                             */
    SGIPC_MAXITER = 3,    /* iteration cap reached before the tolerance           */
    SGIPC_CAPACITY = 4,   /* a caller-provided buffer is too small                 */
    SGIPC_BAD_INDEX = 5,  /* an index lies outside its declared range             */
    SGIPC_BAD_ARG = 6,    /* null pointer, unsupported size or option             */
    SGIPC_HIP_ERROR = 7   /* a HIP runtime call failed                            */
};

enum sgipc_precision {
    SGIPC_FP64 = 0,       /* FP64 SpMV and FP64 MAS apply (the plain path)                */
    SGIPC_FP32_MAS = 1,   /* FP64 SpMV/vectors, FP32 MAS apply (the reference's setting)  */
    SGIPC_FP32_IR = 2     /* FP32 SpMV + FP32 MAS inside FP64 iterative refinement        */
};

enum sgipc_stop_mode {
    SGIPC_STOP_REF_RZ = 0,        /* |r.z| <= tol * r0.z0 (pcg_solver.cu:279-280)       */
    SGIPC_STOP_TRUE_RESIDUAL = 1  /* ||b - A x|| <= tol * ||b||                         */
};

typedef struct sgipc_ctx sgipc_ctx;
typedef struct sgipc_mas sgipc_mas;
typedef struct sgipc_system sgipc_system;
typedef struct sgipc_spmat sgipc_spmat;

/* ---- context ------------------------------------------------------------- */
int sgipc_abi_version(void);
int sgipc_create(sgipc_ctx** out, int device);
int sgipc_destroy(sgipc_ctx* ctx);
/* Use an existing HIP stream (hipStream_t, host handle) instead of the context's own. */
int sgipc_set_stream(sgipc_ctx* ctx, void* hip_stream);
int sgipc_synchronize(sgipc_ctx* ctx);
const char* sgipc_last_error(const sgipc_ctx* ctx);

/* ---- device memory helpers (for bindings without a device allocator) ----- */
int sgipc_malloc(sgipc_ctx* ctx, void** ptr, size_t bytes);
int sgipc_free(sgipc_ctx* ctx, void* ptr);
int sgipc_memcpy_htod(sgipc_ctx* ctx, void* dst, const void* host_src, size_t bytes);
int sgipc_memcpy_dtoh(sgipc_ctx* ctx, void* host_dst, const void* src, size_t bytes);
int sgipc_memcpy_dtod(sgipc_ctx* ctx, void* dst, const void* src, size_t bytes);
int sgipc_memset(sgipc_ctx* ctx, void* dst, int value, size_t bytes);

/* ---- Hessian assembly (triplet writers) ------------------------------------
 * Replaces _calculate_fem_gradient_hessian (StiffGIPC/femEnergy.cu:1719-1857):
 * stable Neo-Hookean tets, gradient += dt^2 vol P:dF/dx (grad may be NULL),
 * PSD-projected Hessian as 10 upper blocks per tet written at
 * tvals/trows/tcols[10*t ...].  tets: uint4 per tet; dm_inv: row-major 3x3;
 * mu/lam: per-tet lengthRate/volumeRate; hess_offset is added to vertex ids. */
int sgipc_fem_tet_gradient_hessian(sgipc_ctx* ctx, const double* x, const uint32_t* tets,
                                   const double* dm_inv, const double* vol, const double* mu,
                                   const double* lam, int n_tets, double dt, int hess_offset,
                                   double* grad, double* tvals, int* trows, int* tcols);

/* Replaces the point-triangle branch of _calBarrierGradientAndHessian
 * (StiffGIPC/GIPC.cu:4733-4984, RANK 2, NEWF) together with the contact
 * lower-block swap (StiffGIPC/GIPC.cu:10329-10363): 16 blocks per pair at
 * tvals[16*p ...], blocks below the diagonal moved to (col,row) and zeroed.
 * dhat_sq is the reference's dHat (squared distance).  pair_x (optional, 12
 * doubles per pair) supplies the pair's local geometry; NULL reads x at the
 * pair's vertex ids like the reference. */
int sgipc_barrier_pt_gradient_hessian(sgipc_ctx* ctx, const double* x, const int* pairs,
                                      const double* pair_x, int n_pairs, double dhat_sq,
                                      double kappa, int hess_offset, double* grad, double* tvals,
                                      int* trows, int* tcols);

/* Replaces lbvh_f/lbvh_e::Construct + SelfCollitionDetect with
 * _checkPTintersection / _checkEEintersection (StiffGIPC/mlbvh.cu:535-1060,
 * 1306-1632, 2084-2226; smooth = false): contact pairs of the surface within
 * squared distance dhat_sq, found by point-face queries (surf_verts against
 * faces, uint3 each) and edge-edge queries (edges, uint2 each) over linear
 * BVHs.  pairs (int4) use the reference's encoding: (-p-1, q, -1, w) point-
 * point, (-p-1, a, b, w) point-edge, (-p-1, t0, t1, t2) point-triangle,
 * (a0, a1, b0, b1) edge-edge, with w = -obj-2 for point-point/point-edge
 * pairs of two edges nearly parallel at rest (rest_x), else -1.  mat_index
 * numbers each class in output order.  The reference's order depends on its
 * atomics; here it is fixed: point-face pairs by (surface slot, face), then
 * edge-edge pairs by (edge, edge).  counts (host, 5 ints) = _cpNum: [0] pairs,
 * [2] point-point, [3] point-edge, [4] point-triangle + edge-edge.  body_id
 * (pairs of one body other than -1 are skipped) and boundary_type (pairs
 * whose four vertices are all >= 2 are skipped) may be NULL.  With too small a
 * capacity: SGIPC_CAPACITY, nothing written, counts[0] = the pairs found. */
int sgipc_contact_detect(sgipc_ctx* ctx, const double* x, const double* rest_x, int n_verts,
                         const uint32_t* faces, int n_faces, const uint32_t* edges, int n_edges,
                         const uint32_t* surf_verts, int n_surf, const int* body_id,
                         const int* boundary_type, double dhat_sq, int* pairs, int* mat_index,
                         int capacity, int* counts);

/* Replaces _calBarrierGradientAndHessian for every pair type (StiffGIPC/
 * GIPC.cu:3176-4984, RANK 2, NEWF: edge-edge, point-point, point-edge,
 * point-triangle) and the contact lower-block swap (GIPC.cu:10344-10362):
 * gradient added atomically into grad (may be NULL), Hessian blocks into the
 * contact arena [PT+EE x 16 | PE x 9 | PP x 4] at mat_index * 16,
 * counts[4] * 16 + mat_index * 9, counts[4] * 16 + counts[3] * 9 +
 * mat_index * 4 (GIPC.cu:3414,4172,4728,4981).  pairs, mat_index and counts
 * as written by sgipc_contact_detect. */
int sgipc_barrier_gradient_hessian(sgipc_ctx* ctx, const double* x, const int* pairs,
                                   const int* mat_index, int n_pairs, const int* counts,
                                   double dhat_sq, double kappa, int hess_offset, double* grad,
                                   double* tvals, int* trows, int* tcols);

/* Replaces lbvh_f/lbvh_e::ConstructFullCCD + SelfCollitionFullDetect and
 * GIPC::self_largestFeasibleStepSize (StiffGIPC/mlbvh.cu:1408-1774,2103-2238;
 * GIPC.cu:7095-7198,9010-9030) with the additive CCD of StiffGIPC/ACCD.cu:
 * 317-534: the largest t in (0, 1] (*step, host) such that x - t dir keeps
 * every candidate pair apart, with eta = 1 - slackness and zero thickness.
 * Candidates: point-face and edge-edge pairs whose boxes, swept to
 * x - alpha dir, overlap within sqrt(dhat_sq) (point-face pairs whose point is
 * behind the face's interior are skipped, as in _checkPTintersection_fullCCD).
 * *n_candidates (host, may be NULL) receives their count. */
int sgipc_ccd_self_step(sgipc_ctx* ctx, const double* x, const double* dir, double alpha,
                        const uint32_t* faces, int n_faces, const uint32_t* edges, int n_edges,
                        const uint32_t* surf_verts, int n_surf, const int* body_id,
                        const int* boundary_type, double dhat_sq, double slackness, double* step,
                        int* n_candidates);

/* Replaces GIPC::ground_largestFeasibleStepSize (StiffGIPC/GIPC.cu:6945-7020,
 * 9070-9090): min over surface vertices moving toward the plane
 * normal . x = offset (normal: device, 3 doubles) of slackness * dist / coef. */
int sgipc_ground_step(sgipc_ctx* ctx, const double* x, const uint32_t* surf_verts, int n_surf,
                      const double* normal, double offset, const double* dir, double slackness,
                      double* step);

/* Replaces GIPC::InjectiveStepSize (StiffGIPC/GIPC.cu:1130-1300,7022-7094,
 * 9091-9111): the smallest t > 0 at which a tet's volume along x - t dir
 * shrinks to slackness times its current volume, minimised over the tets (1
 * when none does).  Below error_rate in the cubic's leading coefficients the
 * reference's quadratic / linear branches; otherwise the cubic's smallest
 * positive root (the reference leaves that case undefined). */
int sgipc_injective_step(sgipc_ctx* ctx, const double* x, const uint32_t* tets, int n_tets,
                         const double* dir, double slackness, double error_rate, double* step);

/* Replaces _calFrictionLastH_DistAndTan (StiffGIPC/GIPC.cu:8431-8535, RANK 2)
 * with the closest-point and tangent-basis helpers of FrictionUtils.cuh:
 * for each pair at the step-start positions x_t, the normal force lambda, the
 * closest-point coordinates dist_coord (2 per pair: PT betas, EE gammas, PE
 * eta, PP zeros) and the tangent basis tan_basis (6 per pair: the two basis
 * vectors, 3 doubles each). */
int sgipc_friction_prepare(sgipc_ctx* ctx, const double* x_t, const int* pairs, int n_pairs,
                           double dhat_sq, double kappa, double* lambda, double* dist_coord,
                           double* tan_basis);

/* Replaces _calFrictionGradient and _calFrictionHessian (StiffGIPC/GIPC.cu:
 * 5130-5313,1482-1845; C1 static-friction clamp, FrictionUtils.cuh:330-500):
 * lagged friction of the step-start pairs (from sgipc_friction_prepare) at x
 * against x_t; eps_sq = the reference's fDhat * dt^2, mu = frictionRate.
 * Gradient added atomically into grad (may be NULL); Hessian blocks
 * [PT+EE x 16 | PE x 9 | PP x 4] by mat_index with counts of those pairs, the
 * lower blocks moved to (col, row) and zeroed. */
int sgipc_friction_gradient_hessian(sgipc_ctx* ctx, const double* x, const double* x_t,
                                    const int* pairs, const int* mat_index, int n_pairs,
                                    const int* counts, const double* lambda,
                                    const double* dist_coord, const double* tan_basis,
                                    double eps_sq, double mu, int hess_offset, double* grad,
                                    double* tvals, int* trows, int* tcols);

/* Replaces _calFrictionLastH_gd (StiffGIPC/GIPC.cu:8405-8429): normal force of
 * each ground vertex at x_t for the plane normal . x = offset (normal: device,
 * 3 doubles, unit). */
int sgipc_ground_friction_prepare(sgipc_ctx* ctx, const double* x_t, const int* verts, int n,
                                  const double* normal, double offset, double dhat_sq,
                                  double kappa, double* lambda);

/* Replaces _calFrictionGradient_gd and _calFrictionHessian_gd (StiffGIPC/
 * GIPC.cu:5090-5128,1400-1480; C0 clamp): one diagonal block per ground vertex
 * at tvals[i]; the tangent projector is I - n n^T (the reference writes the xz
 * entries of its y-up ground, the same values). */
int sgipc_ground_friction_gradient_hessian(sgipc_ctx* ctx, const double* x, const double* x_t,
                                           const int* verts, int n, const double* normal,
                                           const double* lambda, double eps_sq, double mu,
                                           int hess_offset, double* grad, double* tvals,
                                           int* trows, int* tcols);

/* Replaces _calculate_triangle_fem_gradient_hessian (StiffGIPC/femEnergy.cu:
 * 2146-2215): Baraff-Witkin stretch + shear cloth triangles, gradient
 * (atomically added into grad when non-NULL) and projected Hessian as 6 upper
 * blocks per triangle at tvals/trows/tcols[6*t ...].  tris: uint3 per
 * triangle; dm_inv2: 2x2 rest inverse, row-major; area: rest area times
 * thickness. */
int sgipc_fem_tri_gradient_hessian(sgipc_ctx* ctx, const double* x, const uint32_t* tris,
                                   const double* dm_inv2, const double* area, int n_tris,
                                   double k_stretch, double k_shear, double strain_rate, double dt,
                                   int hess_offset, double* grad, double* tvals, int* trows,
                                   int* tcols);

/* Replaces PrepareQuadBendingQ (StiffGIPC/femEnergy.cu:2490-2618, a host loop
 * there): Q (16 doubles, row-major 4x4) per hinge from rest positions.
 * edges, adj: uint2 per hinge; adj.y == 0xFFFFFFFF marks a boundary edge. */
int sgipc_quad_bending_prepare(sgipc_ctx* ctx, const double* rest_x, const uint32_t* edges,
                               const uint32_t* adj, int n_edges, double* q16);

/* Replaces _calculate_quad_bending_gradient_hessian (StiffGIPC/femEnergy.cu:
 * 2620-2754): gradient coef*Q*x and 10 upper blocks coef*Q_ab*I per hinge
 * (coef = |rest edge| * k_bend * dt^2); boundary hinges write ten zero blocks
 * at (0, 0). */
int sgipc_quad_bending_gradient_hessian(sgipc_ctx* ctx, const double* x, const double* rest_x,
                                        const uint32_t* edges, const uint32_t* adj,
                                        const double* q16, int n_edges, double k_bend, double dt,
                                        int hess_offset, double* grad, double* tvals, int* trows,
                                        int* tcols);

/* Replaces ABDSystem::_cal_abd_body_gradient_and_hessian and the body part of
 * _setup_abd_system_hessian (StiffGIPC/abd_system/abd_system_function/
 * setup_abd_system_gradient_and_hessian.cu:200-339,464-476): per body
 * H = M + PSD(kappa vol dt^2 d2V/da2), G = M (q - q~) + kappa vol dt^2 dV/da,
 * 10 upper blocks at rows 4b..4b+3 written to tvals[10*b ...]; grad
 * (optional) receives 12 per body.  body_fixed[b] (NULL: all free) is the
 * reference's BodyBoundaryType: 0 free, 1 fixed (G = 0, H = M), 2 motor
 * (G += W (q - q_p), H += W on the a2, a3 rows, W = motor_strength times the
 * body's 6x6 mass block there, q_p = q~ turned by motor_speed * dt about the
 * body x axis; lines 271-339). */
int sgipc_abd_body_gradient_hessian(sgipc_ctx* ctx, int n_bodies, const double* q,
                                    const double* q_tilde, const double* mass10,
                                    const double* body_volume, double kappa, double dt,
                                    const int* body_fixed, double motor_speed, double motor_strength,
                                    double* grad, double* tvals, int* trows, int* tcols);

/* Replaces GIPC::partitionContactHessian with the ABD contact transforms
 * (StiffGIPC/GIPC.cu:10080-10256; setup_abd_system_gradient_and_hessian.cu:
 * 87-141,424-558): point-level upper contact blocks become, class by class in
 * input order, FEM-FEM blocks (rows shifted by 4*n_bodies - n_abd_points),
 * 4 ABD-FEM blocks J^T H and 16 ABD-ABD blocks J_i^T H J_j (lower ones
 * zeroed).  Blocks of a fixed body (body_fixed == 1) and blocks with a FEM
 * end whose fem_boundary_type (per FEM vertex, NULL: none) is nonzero are
 * written as zeros (GIPC.cu:10332-10363; setup_abd_system_gradient_and_
 * hessian.cu:564-568).  *out_n receives the count; SGIPC_CAPACITY if it
 * exceeds capacity (nothing written then). */
int sgipc_contact_partition(sgipc_ctx* ctx, const double* vals, const int* rows, const int* cols,
                            int n, int n_abd_points, const int* point_body, const double* x_bar,
                            int n_bodies, const int* body_fixed, const int* fem_boundary_type,
                            double* ovals, int* orows, int* ocols, int capacity, int* out_n);

/* Replaces ABDSystem::_cal_abd_system_preconditioner (setup_abd_system_
 * gradient_and_hessian.cu:560-655): inverse of every body's 12x12 diagonal
 * block of the assembled upper BCOO (row-major, 144 doubles per body);
 * SGIPC_NOT_SPD on a singular block.  apply: z = P r on the first
 * 12*n_bodies entries. */
int sgipc_abd_jacobi_build(sgipc_ctx* ctx, const double* vals, const int* rows, const int* cols,
                           int nnz, int n_bodies, double* inv144);
int sgipc_abd_jacobi_apply(sgipc_ctx* ctx, const double* inv144, int n_bodies, const double* r,
                           double* z);

/* Replaces the mass-diagonal writer of computeGradientAndHessian
 * (StiffGIPC/GIPC.cu:10466-10485): mass[i]*I at (i+off, i+off). */
int sgipc_mass_diagonal(sgipc_ctx* ctx, const double* mass, int n, int hess_offset,
                        double* tvals, int* trows, int* tcols);

/* Replaces Converter::srbk_convert (StiffGIPC/linear_system/utils/converter.cu:
 * 32-51): radix sort by (row,col), sum duplicates, unique upper BCOO written to
 * ovals/orows/ocols (capacity blocks); *out_nnz receives the count, also
 * with SGIPC_CAPACITY (nothing past capacity is written; call again with at
 * least that many).  The input arrays are left unchanged. */
int sgipc_convert_srbk(sgipc_ctx* ctx, const double* tvals, const int* trows, const int* tcols,
                       int n_triplets, double* ovals, int* orows, int* ocols, int capacity,
                       int* out_nnz);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): sgipc_convert_srbk with the same arguments, results and
 * status codes, run in a fixed per-device scratch array (24 MiB, held by each
 * GPU the library runs on until the process ends) instead of buffers the
 * library allocates; it calls no hipMalloc/hipFree.  At most 262,144 triplets
 * per call (SGIPC_CAPACITY beyond).  Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU (two streams or threads) race. */
int sgipc_convert_srbk_noalloc(sgipc_ctx* ctx, const double* tvals, const int* trows, const int* tcols,
                               int n_triplets, double* ovals, int* orows, int* ocols, int capacity,
                               int* out_nnz);

/* ---- SpMV ------------------------------------------------------------------
 * Replaces Spmv::srbk_spmv (StiffGIPC/linear_system/utils/spmv.cu:188-313):
 * y = a*A*x + b*y for the symmetric matrix stored as upper BCOO. */
int sgipc_spmv_srbk(sgipc_ctx* ctx, double a, const double* vals, const int* rows,
                    const int* cols, int nnz, const double* x, double b, double* y, int n_rows);
/* FP32 variant (matrix, x and y in float). */
int sgipc_spmv_srbk_f32(sgipc_ctx* ctx, float a, const float* vals, const int* rows,
                        const int* cols, int nnz, const float* x, float b, float* y, int n_rows);
/* Round FP64 blocks/vectors to FP32 (n values). */
int sgipc_convert_f64_to_f32(sgipc_ctx* ctx, const double* in, float* out, size_t n);

/* Prepared symmetric matrix for repeated SpMVs (the PCG loop's use of
 * Spmv::srbk_spmv, StiffGIPC/linear_system/solver/pcg_solver.cu:191-332):
 * create() lays the upper BCOO out in row tiles (in-tile blocks stored once,
 * cross-tile blocks in both rows; an FP32 copy when want_f32), apply() runs
 * y = a*A*x + b*y without atomics.  The layout copies the values: the input
 * arrays may change or be freed after create(). */
int sgipc_spmat_create(sgipc_ctx* ctx, sgipc_spmat** out, const double* vals, const int* rows,
                       const int* cols, int nnz, int n_rows, int want_f32);
int sgipc_spmat_destroy(sgipc_spmat* m);
int sgipc_spmat_apply(sgipc_spmat* m, double a, const double* x, double b, double* y);
int sgipc_spmat_apply_f32(sgipc_spmat* m, float a, const float* x, float b, float* y);
/* FP32 copy of the values (want_f32) against FP64 x and y, FP64 accumulation:
 * the SpMV of the FP32_IR solver's inner PCG. */
int sgipc_spmat_apply_f32_values(sgipc_spmat* m, double a, const double* x, double b, double* y);
/* Layout report (no counterpart in the reference): *halo_form = 1 when the
 * column words index each tile's LDS copy of x (own rows + halo), 0 for the
 * global-gather form (taken when that copy does not fit in LDS, or when the
 * environment variable SGIPC_SPMV_GATHER is set to a non-zero value at
 * create()); *halo_max = largest halo of a 512-row tile (vertices);
 * *tl_slots = in-tile transposed slots the fullest tile uses.  Any output
 * pointer may be NULL. */
int sgipc_spmat_layout(const sgipc_spmat* m, int* halo_form, int* halo_max, int* tl_slots);
/* Rows the layout keeps outside the tiles (more than 128 blocks: affine-body
 * rows, hub vertices), multiplied by a kernel of their own one workgroup per
 * row, so that no tile's workgroup carries their length. */
int sgipc_spmat_long_rows(const sgipc_spmat* m, int* n_long);

/* Measurement helper: lays the upper BCOO out as the PCG solver does (row
 * tiles, in-tile blocks once, cross-tile blocks in both rows) and returns the
 * average duration (ms) of the PCG SpMV kernel (emits the p.Ap partials)
 * over `reps` launches, each timed with its own HIP event pair that the
 * launch records at the kernel's start and end (hipExtLaunchKernelGGL), so the
 * figure is the kernel's duration as a kernel trace reports it, without the
 * dispatch gaps of separately recorded events.  cold_cache != 0 streams a
 * 512 MiB buffer before every
 * launch, so the matrix comes from HBM as it does inside PCG, where the MAS
 * apply evicts it; 0 times back-to-back launches.  precision: SGIPC_FP64 or
 * SGIPC_FP32_IR (FP32 matrix values with FP64 vectors, the kernel PCG runs in
 * that mode).  *stored_blocks (optional) receives the number of blocks the
 * layout stores. */
int sgipc_spmv_time(sgipc_ctx* ctx, int precision, const double* vals, const int* rows,
                    const int* cols, int nnz, int n_rows, int reps, int cold_cache, double* avg_ms,
                    int* stored_blocks);

/* ---- MAS preconditioner (MASPreconditioner, StiffGIPC/MASPreconditioner.cuh:59-107) */
/* bank_size is the domain size B (16 = the reference's BANKSIZE, or 32). */
int sgipc_mas_create(sgipc_ctx* ctx, sgipc_mas** out, int bank_size);
int sgipc_mas_destroy(sgipc_mas* mas);
/* initPreconditioner_Neighbor (MASPreconditioner.cu:2291-2328): vertex count,
 * collision node offset and the vertex-neighbour CSR (device arrays, copied). */
int sgipc_mas_init_neighbor(sgipc_mas* mas, int vert_num, int collision_node_offset,
                            int total_neighbor_num, const uint32_t* nbr_start,
                            const uint32_t* nbr_num, const uint32_t* nbr_list);
/* initPreconditioner_Matrix (:2330-2358): sizes the buffers from the
 * contact-free hierarchy. */
int sgipc_mas_init_matrix(sgipc_mas* mas);
/* setPreconditioner_bcoo (:2207-2245): rebuilds the hierarchy with the current
 * contact pairs, assembles the Galerkin domain matrices from the unique upper
 * BCOO (blocks whose row and column both lie in [offset, offset+vert_num)) and
 * inverts them in FP64.  precision: SGIPC_FP64 stores the inverses in FP64,
 * otherwise FP32. */
int sgipc_mas_set_preconditioner_bcoo(sgipc_mas* mas, const double* vals, const int* rows,
                                      const int* cols, int nnz, int offset, const int* pairs,
                                      int n_pairs, int precision);
/* preconditioning (:2248-2289): z = M^-1 r on the vert_num FEM vertices. */
int sgipc_mas_preconditioning(sgipc_mas* mas, const double* r, double* z);
/* Level introspection (host outputs): *level_num, offsets[level_num+1],
 * counts[level_num]. */
int sgipc_mas_levels(sgipc_mas* mas, int* level_num, int* offsets, int* counts);
/* Hierarchy tables (device outputs): coarse_table[vert_num*(level_num-1)]
 * (global slot ids, vertex-major) and going_next[total slots]. */
int sgipc_mas_tables(sgipc_mas* mas, int* coarse_table, int* going_next);
/* Coarse levels of the apply.  0 (default): additive over all levels, as
 * MASPreconditioner::preconditioning (MASPreconditioner.cu:2248-2289).
 * 1: no counterpart in the reference -- the levels >= 1 become a symmetric
 * V(1,1) cycle on level 1 (level-1 domain solves as the smoother, residuals
 * with the level-1 Galerkin operator, and as its coarse correction the exact
 * level-2 solve: a dense FP64 inverse of A_2 = P_12^T A_1 P_12 built by a
 * blocked Gauss-Jordan on the device, up to 4608 unknowns, beyond which the
 * additive levels >= 2 stand in); level 0 stays additive, so the
 * preconditioner is still symmetric.  A level-2 operator whose inverse meets
 * a pivot <= 0 makes sgipc_mas_set_preconditioner_bcoo return SGIPC_NOT_SPD.
 * 2: no counterpart in the reference -- mode 1's level-1 cycle becomes the
 * coarse solve of a symmetric two-grid cycle on level 0: z0 = w M0 r,
 * r1 = r - A z0, c = (level-1 cycle of P^T r1), r2 = r1 - A P c,
 * z = z0 + P c + w M0 r2, with M0 the level-0 domain solves, P the level-1
 * aggregation and w the smoothing weight (default 0.4).  It needs the fine
 * matrix A inside the apply: the PCG solver supplies its own when the MAS
 * covers every row of the system (else the apply falls back to mode 1); a
 * standalone MAS keeps a copy made at set_preconditioner_bcoo.  Symmetric
 * positive definite for weights below 2 / lambda_max(M0 A); PCG stops with
 * SGIPC_NOT_SPD when r.z <= 0 shows that it is not, and the system solver
 * then redoes the solve with half the weight (twice at most and not below
 * 0.05; a weight that cures it is kept for later solves), then with mode 1
 * and the weight restored.  Solve statistics report the path the apply took
 * (sgipc_pcg_stats.coarse_mode), which can be lower than the mode set.
 * Modes take effect at the next sgipc_mas_set_preconditioner_bcoo (or system
 * solve). */
int sgipc_mas_set_coarse_mode(sgipc_mas* mas, int mode);
/* Smoothing weight w of coarse mode 2, 0 < w < 2 (default 0.4; no
 * counterpart in the reference). */
int sgipc_mas_set_smoothing_weight(sgipc_mas* mas, double omega);
/* Storage of the level-0 domain inverses read by the float applies of coarse
 * mode 2 (no counterpart in the reference): on (the default), FP16 scaled to a
 * unit diagonal, C = D^-1 Minv D^-1 with D^2 = diag(Minv) and D kept per slot
 * in float, so that Minv r = D C D r (|C_ij| <= 1 for SPD domain inverses;
 * half the bytes of the float blocks); off, the float blocks.  FP64 applies
 * and the other modes always use the full inverses.  Takes effect at the next
 * sgipc_mas_set_preconditioner_bcoo (or system solve). */
int sgipc_mas_set_level0_half(sgipc_mas* mas, int on);
/* Measurement helper: average duration (ms) of one preconditioning (all of its
 * kernels) over `reps` applications, HIP events on the context's stream. */
int sgipc_mas_time_apply(sgipc_mas* mas, const double* r, double* z, int reps, double* avg_ms);
/* Inverse domain matrices as FP64 upper blocks (device output,
 * total_slots/B * B(B+1)/2 * 9 doubles). */
int sgipc_mas_inverse_blocks(sgipc_mas* mas, double* out);
/* Diagnostics (no counterpart in the reference; the reference's additive MAS
 * has no dense level): the dense inverse of the level-2 operator that coarse
 * modes 1 and 2 apply, as m x m row-major FP64 (device output, NULL: size
 * only); *m = 0 when the build has no exact level-2 solve. */
int sgipc_mas_level2_inverse(sgipc_mas* mas, double* out, int* m);

/* ---- CEMAS partitions (host functions, no device work) ------------------- */
/* Slot maps of a vertex partition (set_mas_partition,
 * app/common/sim_bootstrap.cpp:90-113; the arrays MASPreconditioner reads as
 * d_partId_map_real / d_real_map_partId, MASPreconditioner.cu:2291-2328).
 * part_id[n] (host): part of each vertex, >= 0.  Part p owns slots
 * [bank p, bank p + bank) and its vertices fill them in increasing vertex
 * order; *map_size = (max part + 1) * bank.  partId_map_real[map_size] gets
 * the vertex of each slot (-1: empty), real_map_partId[n] the slot of each
 * vertex (both host; both NULL: size query).  Equal to the reference's arrays
 * for part ids in non-decreasing order (its "_sorted" meshes), which is the
 * only order its loop handles; valid for any order here.
 * SGIPC_CAPACITY: a part with more than bank vertices. */
int sgipc_mas_partition_maps(const int* part_id, int n, int bank, int* map_size, int* partId_map_real,
                             int* real_map_partId);
/* Partition of a vertex graph into parts of 1..bank vertices (replaces the
 * METIS k-way call of MeshProcess/metis_partition/src/metis_sort.cpp:71-110,
 * node_edge_model.cpp:5-28, which retries with more parts until the largest
 * fits a bank): recursive multilevel bisection into ceil(n / (bank - 1))
 * parts, numbered in bisection-tree order (consecutive ids are close, which
 * the coarse MAS levels rely on).  Host CSR as sgipc_mas_init_neighbor takes
 * it (symmetric); part_id[n] (host) out; *n_parts out.  Deterministic. */
int sgipc_partition_vertices(int n, const uint32_t* nbr_start, const uint32_t* nbr_num,
                             const uint32_t* nbr_list, int bank, int* part_id, int* n_parts);

/* ---- PCG (PCGSolver::solve, StiffGIPC/linear_system/solver/pcg_solver.cu:158-332) */
typedef struct sgipc_pcg_opts {
    double tol;        /* relative tolerance (1e-5 for the true-residual rule)        */
    int stop_mode;     /* enum sgipc_stop_mode                                         */
    int max_iter;      /* <= 0: 0.3 * dof like the reference (pcg_solver.cu:174)     */
    int precision;     /* enum sgipc_precision                                         */
    int check_every;   /* iterations between host looks at the stop flag (<= 0: 8)   */
    int ir_max_outer;  /* FP32_IR: refinement steps (<= 0: 20)                         */
    double ir_inner_tol; /* FP32_IR: floor of the inner tolerance (<= 0: none; each pass
                            asks for the residual reduction still missing)           */
    int fuse;          /* 1 (default): fold the x/r update into the first MAS kernel
                          (MAS on every row, >= 3 levels); 2: also fold the p update
                          into the SpMV; 0: separate vector kernels.  Same
                          iterates either way. */
    int flexible;      /* 1: flexible PCG, beta = -(z.Ap)/(p.Ap) (a preconditioner that
                          may change between iterations, e.g. the learned level); needs
                          the MAS on every row.  Same iterates as PCG for a fixed SPD
                          preconditioner, up to rounding. */
    int ir_replace_every; /* FP32_IR: replace the recursive residual by the FP64 one every
                             that many iterations (default 8; 0: none, refinement by
                             restarts only) */
    int graph;         /* 1: the kernels of each full chunk of check_every iterations
                          after the first are captured once per solve into a HIP graph
                          and replayed (fewer launch gaps; same kernels, same iterates);
                          not with flexible PCG or residual recording.  0: eager. */
} sgipc_pcg_opts;

typedef struct sgipc_pcg_stats {
    int iterations;          /* PCG iterations (summed over refinement steps)        */
    int outer_iterations;    /* refinement steps (1 unless FP32_IR)                   */
    double rel_residual;     /* ||b - A x|| / ||b|| in FP64 at exit                   */
    double rz;               /* last r.z                                               */
    double rz0;
    double ms_total;         /* wall time of the call                                  */
    int status;
    int learned;             /* 1: the learned coarse correction was applied; 2: a run  */
                             /* with it did not converge within twice the plain count   */
                             /* (automatic mode) or met r.z <= 0 (any mode), and the    */
                             /* solve was redone without it (iterations count both)     */
    /* system solves (diagnostics; no counterpart in the reference): the MAS    */
    /* coarse mode of the run that returned, whether its level-1 cycle had the  */
    /* exact dense level-2 solve (0: the factorisation was not usable and the    */
    /* additive levels >= 2 stood in), and the mode-2 smoothing weight          */
    int coarse_mode;
    int level2_exact;
    double omega;
    /* 1: a mixed-precision system solve failed (an indefinite FP32/FP16        */
    /* preconditioner, or refinement passes that stopped reducing the FP64      */
    /* residual) and was redone in FP64 with the preconditioner rebuilt in FP64; */
    /* 2: the solve ran in FP64 from the start because mixed precision failed   */
    /* on an earlier solve of the system (it is tried again after 8, 16, 32 ...  */
    /* solves).  Iterations count every run of the call, failed ones included.  */
    int fp64_fallback;
} sgipc_pcg_stats;

void sgipc_pcg_default_opts(sgipc_pcg_opts* opts);
/* x = A^-1 b with A the unique upper BCOO (FP64), preconditioned by mas (NULL =
 * identity), which must have been set on the same matrix.  For SGIPC_FP32_IR
 * the FP32 copy of the matrix is made internally. */
int sgipc_pcg_solve(sgipc_ctx* ctx, const double* vals, const int* rows, const int* cols, int nnz,
                    int n_rows, sgipc_mas* mas, const double* b, double* x,
                    const sgipc_pcg_opts* opts, sgipc_pcg_stats* stats);

/* ---- whole linear system (GlobalLinearSystem::solve_linear_system,
 *      StiffGIPC/linear_system/linear_system/global_linear_system.cu:136-148) --
 * One object owns a scene's device inputs, the triplet arena, the BCOO, the MAS
 * and the PCG buffers; solve() runs assembly in the reference's arena order
 * (contacts, tets, bending hinges, cloth triangles, mass;
 * StiffGIPC/GIPC.cu:10291-10485), conversion, the MAS build and PCG.  Every
 * element set is optional (count 0, pointers NULL). */
typedef struct sgipc_scene_desc {
    int n_verts;
    int n_tets;
    int n_pairs;
    int total_neighbor_num;
    double dt;
    double dhat_sq;
    double kappa;
    /* device arrays */
    const double* x;          /* 3n                                  */
    const uint32_t* tets;     /* 4 per tet                           */
    const double* dm_inv;     /* 9 per tet, row-major                */
    const double* vol;        /* per tet                             */
    const double* mu;         /* per tet, lengthRate                 */
    const double* lam;        /* per tet, volumeRate                 */
    const double* mass;       /* per vertex                          */
    const int* pairs;         /* 4 per pair                          */
    const double* pair_x;     /* 12 per pair or NULL                 */
    const uint32_t* nbr_start;
    const uint32_t* nbr_num;
    const uint32_t* nbr_list;
    /* cloth triangles (Baraff-Witkin) */
    int n_tris;
    double k_stretch, k_shear, strain_rate;
    const uint32_t* tris;     /* 3 per triangle                      */
    const double* dm_inv2;    /* 4 per triangle, row-major           */
    const double* area;       /* per triangle, area * thickness      */
    /* quadratic bending hinges */
    int n_hinges;
    double k_bend;
    const double* rest_x;     /* 3n rest positions                    */
    const uint32_t* hinge_edges; /* 2 per hinge                       */
    const uint32_t* hinge_adj;   /* 2 per hinge, 0xFFFFFFFF = boundary */
    /* affine bodies (ABD).  With n_bodies > 0 the system rows are
     * [4 per body | FEM vertices], contact pairs use global point ids
     * [body points | FEM vertices] and pair_x is required; x, tets, tris,
     * hinges, mass and the neighbour lists stay FEM-local. */
    int n_bodies;
    int n_abd_points;
    double kappa_abd;
    const double* q;          /* 12 per body: p, a1, a2, a3          */
    const double* q_tilde;    /* 12 per body                          */
    const double* mass10;     /* 10 per body: m, m xb, m xb xb^T upper */
    const double* body_volume;
    const int* body_fixed;    /* per body: 0 free, 1 fixed, 2 motor (or NULL) */
    const int* point_body;    /* per body point                       */
    const double* x_bar;      /* 3 per body point, body-frame rest    */
    /* Dirichlet vertices (TetMesh.BoundaryType; GIPC.cu:10332-10363,
     * 10444-10462, fem_linear_subsystem.cu:56-84): per FEM vertex, nonzero =
     * fixed.  Every element and contact block with an end on a fixed vertex
     * is zeroed (its mass block stays) and its gradient rows are zero, so the
     * solve leaves it in place.  NULL: none. */
    const int* boundary_type;
    /* motor bodies (body_fixed == 2): angular speed (rad/s) about the body x
     * axis and strength (multiple of the body mass block); the reference's
     * defaults are 31.4 and 10 (abd_system_parms.h:13-14) */
    double motor_speed;
    double motor_strength;
    /* contact detection on every solve (sgipc_contact_detect, scenes without
     * affine bodies): with detect_contacts = 1, pairs / pair_x / n_pairs are
     * ignored; the pairs of the surface (faces, edges, surf_verts, with
     * rest_x for the nearly-parallel edge test) within dhat_sq give the
     * barrier blocks of every pair type (sgipc_barrier_gradient_hessian) and
     * the MAS contact connectivity, like the reference's Newton step
     * (GIPC.cu:10291-10304).  contact_capacity: pairs held (<= 0: 8 per
     * surface vertex and edge; grown on demand). */
    int detect_contacts;
    int n_faces;
    int n_edges;
    int n_surf;
    const uint32_t* faces;      /* 3 per surface triangle               */
    const uint32_t* edges;      /* 2 per surface edge                   */
    const uint32_t* surf_verts;
    const int* body_id;         /* per vertex or NULL                   */
    int contact_capacity;
    /* CEMAS partition (host array, per FEM vertex, or NULL; FEM-only
     * scenes): the MAS uses the slot maps of sgipc_mas_partition_maps
     * (app/common/sim_bootstrap.cpp:90-113, load_mesh.cpp:447-478) instead of
     * consecutive vertex domains.  The system then solves in slot space: its
     * BCOO (sgipc_system_bcoo) and MAS have one row per slot, the empty slots
     * of a part being rows without blocks; b, x and grad stay per vertex.
     * Parts without a vertex are dropped (they change no MAS level). */
    const int* part_id;
} sgipc_scene_desc;

typedef struct sgipc_solve_timing {
    double ms_assemble;      /* element Hessians + gradient                 */
    double ms_convert;       /* triplets -> unique BCOO                     */
    double ms_mas_build;     /* hierarchy + Galerkin + inverse              */
    double ms_pcg;           /* PCG (incl. refinement)                      */
    double ms_total;
    int nnz;                 /* unique upper blocks                          */
    int n_triplets;
} sgipc_solve_timing;

int sgipc_system_create(sgipc_ctx* ctx, sgipc_system** out, const sgipc_scene_desc* desc,
                        int bank_size);
int sgipc_system_destroy(sgipc_system* sys);
/* b, x: device vectors of 3*n_verts doubles.  grad (optional, device) receives
 * the assembled energy gradient. */
int sgipc_system_solve(sgipc_system* sys, const double* b, double* x, double* grad,
                       const sgipc_pcg_opts* opts, sgipc_pcg_stats* stats,
                       sgipc_solve_timing* timing);
/* Views of the system's current BCOO (device pointers owned by the system;
 * slot rows with a partition). */
int sgipc_system_bcoo(sgipc_system* sys, const double** vals, const int** rows, const int** cols,
                      int* nnz);
sgipc_mas* sgipc_system_mas(sgipc_system* sys);
/* The contact pairs of the last solve (device pointers owned by the system):
 * the detected pairs and their MatIndex with detect_contacts, else the
 * descriptor's pairs (mat_index NULL). */
int sgipc_system_contacts(sgipc_system* sys, const int** pairs, const int** mat_index, int* n_pairs);

/* Replaces GIPC::computeEnergy for the terms the system assembles (StiffGIPC/
 * GIPC.cu:7373-7870): at device positions x, 1/2 |x - x~|^2_M (when x_tilde
 * is non-NULL) + dt^2 (tet, triangle and bending energies) + the barrier of
 * the contact pairs at x (detected there when the scene detects its
 * contacts), summed in a fixed order into *energy (host).  Scenes with affine
 * bodies: SGIPC_BAD_ARG. */
int sgipc_system_energy(sgipc_system* sys, const double* x, const double* x_tilde, double* energy);

/* Replaces GIPC::lineSearch with the feasible-step bound before it (StiffGIPC/
 * GIPC.cu:10828-10920, 9010-9111): from the scene's positions x0 along
 * x0 - a dir, a starts at min(alpha, the CCD self step when contacts are
 * detected, the tet injectivity step) and halves while the energy exceeds
 * the energy at x0 (c1 = 0), at most 8 times; then x0 is restored into x_out
 * and *stopped = 1.  x_out (device, 3 n) receives the accepted positions,
 * *alpha_out the step, energies (host, 2, may be NULL) the energies at x0 and
 * at x_out.  Not included: the reference's edge-triangle intersection
 * re-check and its kappa update after the search. */
int sgipc_system_line_search(sgipc_system* sys, const double* dir, const double* x_tilde, double alpha,
                             double slackness, double* x_out, double* alpha_out, double* energies,
                             int* stopped);

/* ---------------------------------------------------------------------------
 * Learned coarse level.  No counterpart in the reference (it has no learned
 * component): a message-passing network over the MAS level-1 aggregate graph,
 * run with BF16 MFMA on gfx950, whose output is a coarse correction added to
 * the MAS result inside flexible PCG.
 *   X (3 features per row: S P^T r, S^-1 zc, S P^T r - Ahat S^-1 zc)
 *   V0 = X W_in; V <- V + act([V, Ahat V] W_l); chat = V w_out
 * with Ahat = S A_1 S the Jacobi-scaled level-1 Galerkin operator, act = relu
 * (linear == 0) or the identity.  Parameters: W_in (3 x 64), W_l (128 x 64
 * each), w_out (64), row-major, in that order.  Training minimises
 * (0.5 chat^T Ahat chat - chat^T q) / rows per sample with Adam.
 * ------------------------------------------------------------------------- */
typedef struct sgipc_mpnn sgipc_mpnn;
typedef struct sgipc_mpnn_data sgipc_mpnn_data;
int sgipc_mpnn_create(sgipc_ctx* ctx, int layers, int linear, sgipc_mpnn** out);
int sgipc_mpnn_destroy(sgipc_mpnn* net);
int sgipc_mpnn_num_params(const sgipc_mpnn* net);
int sgipc_mpnn_set_params(sgipc_mpnn* net, const float* host, int n);
int sgipc_mpnn_get_params(const sgipc_mpnn* net, float* host, int n);
int sgipc_mpnn_init_params(sgipc_mpnn* net, unsigned long long seed);
/* A training set over one level-1 graph, given as the upper BCOO of A_1
 * (device pointers, FP64 column-major blocks, sorted unique).  A training
 * set, like a network, uses its context's device and stream: destroy it
 * before that context. */
int sgipc_mpnn_data_create(sgipc_ctx* ctx, int n_nodes, const double* vals, const int* rows,
                           const int* cols, int nb, sgipc_mpnn_data** out);
/* Append samples: X [n_samples][3 n_nodes][3], q [n_samples][3 n_nodes] (device). */
int sgipc_mpnn_data_add(sgipc_mpnn_data* d, const float* X, const float* q, int n_samples);
int sgipc_mpnn_data_size(const sgipc_mpnn_data* d);
int sgipc_mpnn_data_nodes(const sgipc_mpnn_data* d);
int sgipc_mpnn_data_samples(const sgipc_mpnn_data* d, int first, int count, float* X_host,
                            float* q_host);
/* The graph as built: CSR over nodes (both directions, diagonal included),
 * Jacobi-scaled row-major 3x3 blocks, node scales (host buffers). */
int sgipc_mpnn_data_graph(const sgipc_mpnn_data* d, int* row_ptr, int* col, float* blk, float* scale);
int sgipc_mpnn_data_graph_nnz(const sgipc_mpnn_data* d);
int sgipc_mpnn_data_destroy(sgipc_mpnn_data* d);
/* chat of samples [first, first+count) into chat (device, count * 3 n_nodes). */
int sgipc_mpnn_forward(sgipc_mpnn* net, const sgipc_mpnn_data* d, int first, int count, float* chat);
/* Mean loss and gradient (host, num_params floats) over samples [first, first+count). */
int sgipc_mpnn_loss_grad(sgipc_mpnn* net, const sgipc_mpnn_data* d, int first, int count,
                         double* loss, float* grad_host);
/* Adam over shuffled minibatches; epoch_loss (host, optional) gets the mean
 * minibatch loss of every epoch. */
int sgipc_mpnn_train(sgipc_mpnn* net, sgipc_mpnn_data* d, int epochs, int batch, float lr,
                     unsigned seed, float* epoch_loss);
/* Training samples from a system: one plain solve of b (learned level off)
 * records the residual every `stride` iterations (at most max_samples); each
 * becomes (X, q = rho) over the system's level-1 graph, normalised by the RMS
 * of ghat.  FEM-only systems.  The set belongs to the system's context
 * (destroy it before that context). */
int sgipc_system_collect(sgipc_system* sys, const double* b, const sgipc_pcg_opts* opts, int stride,
                         int max_samples, sgipc_mpnn_data** out);
/* Attach a network to a system's MAS.  mode 0: off; 1: always (flexible PCG);
 * 2: automatic -- a solve without and one with the correction, then the one
 * with fewer iterations (falls back to plain MAS whenever the learned
 * iteration count is worse); the other is measured again 8 solves later, the
 * gap doubling each time the learned level loses.  A run with the correction
 * is capped at twice the plain count plus 16 iterations: when it does not
 * converge there it loses and the same call solves again without it.
 * stats.learned reports which one ran. */
int sgipc_system_set_learned(sgipc_system* sys, sgipc_mpnn* net, int mode);

#ifdef __cplusplus
}
#endif
#endif /* SGIPC_H */
