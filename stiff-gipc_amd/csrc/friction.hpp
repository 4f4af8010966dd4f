// Lagged IPC friction launchers (pairs and ground).
#pragma once
#include "sgipc_internal.hpp"

namespace sgipc {

// Per pair at the step start xt: normal force, closest-point coordinates (2)
// and tangent basis (6).  A smooth-encoded pair sets *bad.
void launch_friction_prepare(hipStream_t s, int n, const double* xt, const int* pairs, double dhat_sq, double kappa,
                             double* lambda, double* coord, double* basis, int* bad);
// Gradient (atomically added into grad when non-null) and the Hessian blocks
// [PT+EE x 16 | PE x 9 | PP x 4] by MatIndex.
void launch_friction(hipStream_t s, int n, const double* x, const double* xt, const int* pairs, const int* mat_index,
                     int n4, int n3, const double* lambda, const double* coord, const double* basis, double eps2,
                     double mu, int hess_offset, double* grad, double* tvals, int* trows, int* tcols, int* bad);
void launch_ground_friction_prepare(hipStream_t s, int n, const double* xt, const int* verts, const double* normal,
                                    double offset, double dhat_sq, double kappa, double* lambda);
// One diagonal block per ground vertex at tvals[i].
void launch_ground_friction(hipStream_t s, int n, const double* x, const double* xt, const int* verts,
                            const double* normal, const double* lambda, double eps2, double mu, int hess_offset,
                            double* grad, double* tvals, int* trows, int* tcols);

}  // namespace sgipc
