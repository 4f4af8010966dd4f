/*
 * CPU ORACLE for the Newton linear-solve path -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's algorithms, used by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py as the checker
 * (and as the CPU baseline "port").  The product library never links, loads or
 * calls it.
 *
 * Conventions shared with the product ABI:
 *   - a 3x3 block is 9 doubles, column-major (Eigen::Matrix3d): v[c*3 + r] = M(r, c);
 *   - vectors are contiguous doubles of length 3*rows (double3 aliasing);
 *   - the assembled matrix is upper-triangular symmetric BCOO (row <= col),
 *     sorted by (row, col), unique.
 *
 * Pinning (what each function is checked against, see tests/test_oracle_*.py):
 *   orc_tet_snh       energy closed form -> finite-difference gradient; analytic
 *                     unprojected Hessian -> FD of the gradient; projected
 *                     Hessian == numpy eigh-clamp; rest-state eigenvalues.
 *   orc_tri_bw        energy closed form -> FD gradient; each projected term ==
 *                     numpy eigh-clamp of the FD Hessian of dE/dF (stretch-only
 *                     and shear-only), additivity in (k_stretch, k_shear).
 *   orc_quad_bending  Q 1 = 0, rank 1, PSD; equilateral hinge closed form; FD
 *                     gradient of the quadratic energy; flat rest -> zero force.
 *   orc_barrier_pt    distance closed form -> FD gradient; lambda0 == d2E/ds2 (FD);
 *                     rank-1 and PSD.
 *   orc_convert       numpy lexsort + add.reduceat.
 *   orc_spmv          scipy.sparse product of A_upper + A_upper^T - diag; dense.
 *   orc_mas_*         independent set-based restatement in numpy/scipy (per-domain
 *                     connected components, Galerkin P^T A P, numpy.linalg.inv,
 *                     dense apply); M^-1 symmetric positive definite on small meshes.
 *   orc_pcg           scipy.sparse.linalg.cg with the same preconditioner (iterates),
 *                     true residual.
 */
#ifndef SGIPC_ORACLE_H
#define SGIPC_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* ---- element Hessians ------------------------------------------------------ */
/* Stable Neo-Hookean (variant 1) tet: gradient (atomically accumulated into
 * grad when non-NULL) and PSD-projected Hessian written as 10 upper blocks per
 * tet at tvals/trows/tcols[10*t ...].  StiffGIPC/femEnergy.cu:1719-1857. */
int orc_tet_snh(int n_tets, const double* x, const int* tets, const double* dm_inv,
                const double* vol, const double* mu, const double* lam, double dt,
                int hess_offset, double* grad, double* tvals, int* trows, int* tcols);

/* Stable Neo-Hookean energy of one tet (for finite-difference checks). */
double orc_tet_snh_energy(const double* x12, const double* dm_inv, double mu, double lam);
/* Unprojected 9x9 dPsi/dF^2 (column-major vec(F)) and the projected one. */
void orc_snh_dpdf(const double* F9, double mu, double lam, double* H81, double* Hproj81);

/* Point-triangle barrier (GIPC rank-2 variant), 16 blocks per pair at
 * tvals[16*p ...]; blocks below the diagonal are moved to (col,row) and zeroed,
 * as the reference's contact partition does.  StiffGIPC/GIPC.cu:4733-4984,
 * 10329-10363.  pair_x: 12 doubles of local geometry per pair, or NULL to read
 * x at the pair's vertex ids.  dhat_sq is the reference's dHat (a squared
 * distance). */
int orc_barrier_pt(int n_pairs, const double* x, const int* pairs, const double* pair_x,
                   double dhat_sq, double kappa, int hess_offset, double* grad,
                   double* tvals, int* trows, int* tcols);
double orc_barrier_energy(double I5, double dhat_sq, double kappa);
double orc_barrier_lambda0(double I5, double dhat_sq, double kappa);

/* Contact pairs within squared distance dhat_sq (contact_oracle.c): the
 * reference's point-face and edge-edge self queries (StiffGIPC/mlbvh.cu:
 * 535-1060,1306-1632) by brute force, pairs in the reference's int4 encoding
 * in a fixed order (point-face by (surface slot, face), then edge-edge by
 * (edge, edge)), MatIndex per class; counts[0] total, counts[2..4] the PP,
 * PE and PT+EE class sizes (like _cpNum).  Returns -1 when capacity is
 * short (counts[0] holds the total then). */
int orc_contact_detect(int n_verts, const double* x, const double* rest, int n_faces, const int* faces, int n_edges,
                       const int* edges, int n_surf, const int* surf, const int* body_id, const int* btype,
                       double dhat_sq, int capacity, int* pairs, int* mat_index, int* counts);
/* Barrier gradient and Hessian of PP, PE, PT and EE pairs (GIPC.cu:
 * 3176-4984, RANK 2, NEWF) into the contact arena [PT+EE x 16 | PE x 9 |
 * PP x 4] by MatIndex; n4, n3 the PT+EE and PE class sizes. */
int orc_barrier(int n_pairs, const double* x, const int* pairs, const int* mat_index, int n4, int n3, double dhat_sq,
                double kappa, int hess_offset, double* grad, double* tvals, int* trows, int* tcols);
double orc_barrier_energy_pairs(int n_pairs, const double* x, const int* pairs, double dhat_sq, double kappa);

/* Lagged IPC friction (friction_oracle.c; StiffGIPC/GIPC.cu:1400-1845,
 * 5090-5313, 8405-8535; FrictionUtils.cuh): per-pair normal force, closest-
 * point coordinates (2) and tangent basis (6: two 3-vectors) at the step
 * start xt; gradient and Hessian blocks [PT+EE x 16 | PE x 9 | PP x 4] by
 * MatIndex; the ground's per-vertex force and diagonal blocks; energies for
 * the finite-difference pins. */
/* Feasible step sizes (contact_oracle.c): additive CCD over the swept-box
 * candidates (StiffGIPC/ACCD.cu:317-534, mlbvh.cu:1408-1774, GIPC.cu:
 * 7095-7198), the ground step (GIPC.cu:6945-7020) and the tet injectivity
 * step (GIPC.cu:1130-1300,7022-7094); positions move as x - t dir. */
double orc_ccd_self_step(int n_verts, const double* x, const double* dir, double alpha, int n_faces, const int* faces,
                         int n_edges, const int* edges, int n_surf, const int* surf, const int* body_id,
                         const int* btype, double dhat_sq, double slackness, int* n_candidates);
double orc_ground_step(int n_surf, const double* x, const int* surf, const double* normal, double offset,
                       const double* dir, double slackness);
double orc_injective_step(int n_tets, const double* x, const int* tets, const double* dir, double slackness,
                          double error_rate);
int orc_friction_prepare(int n, const double* xt, const int* pairs, double dhat_sq, double kappa, double* lambda,
                         double* coord, double* basis);
int orc_friction(int n, const double* x, const double* xt, const int* pairs, const int* mat_index, int n4, int n3,
                 const double* lambda, const double* coord, const double* basis, double eps2, double mu, int hess_offset,
                 double* grad, double* tvals, int* trows, int* tcols);
double orc_friction_energy(int n, const double* x, const double* xt, const int* pairs, const double* lambda,
                           const double* coord, const double* basis, double eps2, double mu);
int orc_ground_friction_prepare(int n, const double* xt, const int* verts, const double* normal, double offset,
                                double dhat_sq, double kappa, double* lambda);
int orc_ground_friction(int n, const double* x, const double* xt, const int* verts, const double* normal,
                        const double* lambda, double eps2, double mu, int hess_offset, double* grad, double* tvals,
                        int* trows, int* tcols);
double orc_ground_friction_energy(int n, const double* x, const double* xt, const int* verts, const double* normal,
                                  const double* lambda, double eps2, double mu);

/* Baraff-Witkin cloth triangle (stretch + shear), 6 upper blocks per triangle
 * at tvals[6*t ...]; dm_inv2 is the 2x2 rest inverse (row-major), area the
 * rest area.  StiffGIPC/femEnergy.cu:2146-2215. */
int orc_tri_bw(int n_tris, const double* x, const int* tris, const double* dm_inv2, const double* area,
               double k_stretch, double k_shear, double strain_rate, double dt, int hess_offset,
               double* grad, double* tvals, int* trows, int* tcols);
double orc_tri_bw_energy(const double* x9, const double* dm_inv2, double area, double k_stretch,
                         double k_shear, double strain_rate);
/* dE/dF (P6) and projected d2E/dF2 (6x6 row-major), vec(F) column-major, per unit area. */
void orc_bw_dpdf(const double* F6, double k_stretch, double k_shear, double strain_rate, double* P6,
                 double* Hproj36);

/* Quadratic bending: Q (4x4 row-major) per hinge from rest positions
 * (femEnergy.cu:2490-2618), and the gradient / 10 upper blocks per hinge
 * (femEnergy.cu:2620-2754).  edges, adj: 2 ints per hinge; adj[1] == -1 marks
 * a boundary edge. */
void orc_quad_bending_q(int n_edges, const double* rest_x, const int* edges, const int* adj, double* Q16);
int orc_quad_bending(int n_edges, const double* x, const double* rest_x, const int* edges, const int* adj,
                     const double* Q16, double k_bend, double dt, int hess_offset, double* grad,
                     double* tvals, int* trows, int* tcols);

/* mass[i] * I at (i+off, i+off).  StiffGIPC/GIPC.cu:10466-10485. */
int orc_mass_diag(int n, const double* mass, int hess_offset, double* tvals, int* trows,
                  int* tcols);

/* ---- triplets -> unique upper BCOO (Converter::srbk_convert) ---------------- */
long orc_convert(long n_trip, const double* tvals, const int* trows, const int* tcols,
                 double* ovals, int* orows, int* ocols);

/* ---- symmetric upper-BCOO SpMV: y = a*A*x + b*y (Spmv::srbk_spmv) --------- */
void orc_spmv(double a, long nnz, const double* vals, const int* rows, const int* cols,
              const double* x, double b, double* y, int n_rows);

/* ---- MAS preconditioner (StiffGIPC/MASPreconditioner.cu) --------------------- */
typedef struct orc_mas orc_mas;
orc_mas* orc_mas_create(int n, int bank, const unsigned* nbr_start, const unsigned* nbr_num,
                        const unsigned* nbr_list, int collision_node_offset);
void orc_mas_destroy(orc_mas* m);
int orc_mas_build(orc_mas* m, long nnz, const double* vals, const int* rows, const int* cols,
                  int offset, const int* pairs, int n_pairs);
void orc_mas_apply(orc_mas* m, const double* r, double* z, int fp32);
/* level_num; offsets[level_num+1]; counts[level_num]. */
int orc_mas_levels(const orc_mas* m, int* level_num, int* offsets, int* counts);
/* coarse_table[n*(L-1)] (global slot ids, vertex-major), going_next[total slots]. */
void orc_mas_tables(const orc_mas* m, int* coarse_table, int* going_next);
/* Domain matrices, B(B+1)/2 blocks of 9 doubles per domain (inverse=0: Galerkin
 * blocks, inverse=1: inverse blocks). */
void orc_mas_matrices(const orc_mas* m, double* out, int inverse);
int orc_mas_num_levels_for(int n, int bank);

/* ---- PCG (PCGSolver::pcg) ------------------------------------------------- */
/* stop_mode 0: reference |r.z| <= tol * r0.z0 checked after the x/r update
 *              with the previous rz (pcg_solver.cu:279-280);
 * stop_mode 1: recursive ||r||/||b|| <= tol.
 * mas == NULL -> identity preconditioner.  stats[0]=iterations performed
 * (x/r updates; the reference returns max_iter when the cap is hit),
 * stats[1]=true relative residual, stats[2]=final rz, stats[3]=rz0. */
int orc_pcg(long nnz, const double* vals, const int* rows, const int* cols, int n_rows,
            orc_mas* mas, const double* b, double* x, double tol, int stop_mode, int max_iter,
            int mas_fp32, double* stats);

/* ---- affine bodies (ABD) ---------------------------------------------------
 * Body term: 10 upper blocks per body at rows 4b..4b+3 (kinetic M(q - q~) and
 * the PSD-projected orthogonality energy kappa*vol*dt^2 V(a)); grad (optional)
 * receives 12 per body.  fixed[b]: 0 free, 1 fixed (H = M, G = 0), 2 motor
 * (the motor term of lines 271-339 added).
 * setup_abd_system_gradient_and_hessian.cu:200-339. */
int orc_abd_body(int n_bodies, const double* q, const double* q_tilde, const double* mass10,
                 const double* volume, double kappa, double dt, const int* fixed, double motor_speed,
                 double motor_strength, double* grad, double* tvals, int* trows, int* tcols);
double orc_abd_shape_energy(const double* q12);
void orc_abd_shape_grad_hess(const double* q12, double* g9, double* H81);
/* Vertex-level contact triplets (upper, as the barrier writes them) -> system
 * triplets by class, in the order FEM-FEM (1 block), ABD-FEM (4 blocks J^T H),
 * ABD-ABD (16 blocks J_i^T H J_j, lower ones zeroed at (col,row); a same-body
 * pair i != j stands for both (i,j) and (j,i)).  Point ids < n_abd_points are
 * ABD points; FEM point j goes to row 4*n_bodies + j - n_abd_points.  Returns
 * the count.  GIPC.cu:10080-10256, setup_abd_system_gradient_and_hessian.cu:
 * 87-141,424-558. */
long orc_abd_contact_map(long n_trip, const double* vals, const int* rows, const int* cols,
                         int n_abd_points, const int* point_body, const double* x_bar, int n_bodies,
                         const int* body_fixed, const int* fem_btype, double* ovals, int* orows,
                         int* ocols);
/* Dirichlet vertices (GIPC.cu:10444-10462): zero every triplet with an end on
 * a FEM row (>= row0) whose vertex btype[row - row0] is nonzero. */
void orc_dirichlet_zero(long n_trip, const int* rows, const int* cols, int row0, const int* btype,
                        double* vals);
/* Inverse of every body's 12x12 diagonal block of the assembled BCOO
 * (row-major, 144 per body).  setup_abd_system_gradient_and_hessian.cu:560-655. */
int orc_abd_jacobi_build(long nnz, const double* vals, const int* rows, const int* cols, int n_bodies,
                         double* inv144);

/* PCG with the subsystem preconditioner: z = r, then block-Jacobi on the
 * body rows [0, 4*n_bodies) when abd_inv is set and MAS on the rows from
 * mas_row_offset when mas is set (GlobalLinearSystem::apply_preconditioner,
 * global_linear_system.cu:150-220). */
typedef struct orc_precond {
    orc_mas* mas;
    int mas_row_offset;
    int mas_fp32;
    int n_bodies;
    const double* abd_inv;
} orc_precond;
int orc_pcg2(long nnz, const double* vals, const int* rows, const int* cols, int n_rows,
             const orc_precond* pc, const double* b, double* x, double tol, int stop_mode, int max_iter,
             double* stats);

/* Deliberate-mistake switch for the tests that show the checks are sensitive:
 * 0 = none, 1 = SpMV drops the transposed (lower) contribution, 2 = Galerkin
 * coarse levels drop H^T on merged diagonals, 3 = barrier Hessian uses lambda0
 * with the wrong sign of the log term, 4 = cloth shear Hessian left unprojected. */
/* Energy sums of the Newton objective (line search), scaled like the
 * assembled gradients: tets dt^2 vol Psi, triangles dt^2 area Psi, bending
 * 1/2 |e| k_bend dt^2 x^T Q x. */
double orc_energy_tets(int n, const double* x, const int* tets, const double* dm_inv, const double* vol,
                       const double* mu, const double* lam, double dt);
double orc_energy_tris(int n, const double* x, const int* tris, const double* dm_inv2, const double* area,
                       double k_st, double k_sh, double rate, double dt);
double orc_energy_bending(int n, const double* x, const double* rest_x, const int* edges, const int* adj,
                          const double* Q16, double k_bend, double dt);
void orc_set_fault(int fault);
/* OpenMP threads of the parallel loops (k > 0). */
void orc_set_threads(int k);
int orc_get_fault(void);

#ifdef __cplusplus
}
#endif
#endif
