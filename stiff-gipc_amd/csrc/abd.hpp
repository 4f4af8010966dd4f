// Affine-body (ABD) terms, contact class mapping and the body block-Jacobi.
#pragma once
#include "pcg.hpp"
#include "sgipc_internal.hpp"

namespace sgipc {

// 10 upper blocks per body at rows 4b..4b+3; grad (optional) 12 per body.
// fixed[b]: 0 free, 1 fixed, 2 motor (BodyBoundaryType)
void launch_abd_body(hipStream_t s, int nb, const double* q, const double* q_tilde, const double* mass10,
                     const double* vol, double kappa, double dt, const int* fixed, double motor_speed,
                     double motor_strength, double* grad, double* tvals, int* trows, int* tcols);
// point-space energy gradient gp (3 per point: body points, then FEM points)
// added into the system gradient (12 per body, then 3 per FEM vertex)
void launch_point_gradient(hipStream_t s, int n_points, int n_abd, const double* gp, const int* point_body,
                           const double* x_bar, const int* body_fixed, int nb, double* grad);
// zero the triplets touching a Dirichlet FEM vertex (rows >= row0 are FEM rows)
void launch_dirichlet_zero(hipStream_t s, long n, const int* rows, const int* cols, int row0, const int* btype,
                           double* vals);
// zero the gradient rows of Dirichlet vertices (g: 3 per FEM vertex)
void launch_dirichlet_zero_rows(hipStream_t s, int n, const int* btype, double* g);

// Point-level upper contact blocks -> system blocks by class (FEM-FEM,
// ABD-FEM, ABD-ABD), each class in input order.  Returns the block count.
class ContactMapper {
  public:
    // classify the n input blocks; returns the output block count (reads it back)
    long count(hipStream_t s, const int* rows, const int* cols, long n, int n_abd);
    // write the blocks counted by the last count() (same rows/cols)
    void write(hipStream_t s, const double* vals, const int* rows, const int* cols, int n_abd, const int* point_body,
               const double* x_bar, int nb, const int* body_fixed, const int* fem_btype, double* ov, int* orow,
               int* ocol,
               long capacity);

  private:
    DevBuf<int> cnt_[3], scan_[3];
    DevBuf<unsigned char> temp_;
    long n_ = 0, total_ = 0;
};

// Inverse 12x12 diagonal blocks (144 per body) of the body rows [0, 4 nb);
// *bad is or-ed with 1 on a singular block.
void launch_abd_jacobi_build(hipStream_t s, const double* vals, const int* rows, const int* cols, int nnz, int nb,
                             double* inv, int* bad);
// z = P r on the first 12 nb entries, r.z partials; nothing when ctl->done.
template <typename TV>
int launch_abd_jacobi_apply(hipStream_t s, const double* inv, int nb, const TV* r, TV* z, double* rz_part,
                            const PcgCtl* ctl);

// Inverse 12x12 diagonal blocks of the body rows [0, 4 nb).
class AbdJacobi {
  public:
    void build(hipStream_t s, const double* vals, const int* rows, const int* cols, int nnz, int nb);
    // z = P r on the first 12 nb entries, r.z partials; nothing when ctl->done.
    template <typename TV>
    int apply(hipStream_t s, const TV* r, TV* z, double* rz_part, const PcgCtl* ctl) const;
    int partials() const;

    int n_bodies = 0;
    DevBuf<double> inv;

  private:
    DevBuf<int> bad_;
};

}  // namespace sgipc
