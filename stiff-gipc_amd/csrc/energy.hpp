// Energy terms of the Newton objective (line search), summed in a fixed order.
#pragma once
#include <algorithm>

#include "sgipc_internal.hpp"

namespace sgipc {

class EnergySum {
  public:
    double inertia(hipStream_t s, int n, const double* x, const double* x_tilde, const double* mass);
    double tets(hipStream_t s, int n, const double* x, const uint32_t* tets, const double* dm_inv, const double* vol,
                const double* mu, const double* lam, double dt);
    double tris(hipStream_t s, int n, const double* x, const uint32_t* tris, const double* dm_inv2, const double* area,
                double k_stretch, double k_shear, double strain_rate, double dt);
    double bending(hipStream_t s, int n, const double* x, const double* rest, const uint32_t* edges,
                   const uint32_t* adj, const double* q16, double k_bend, double dt);
    double barrier(hipStream_t s, int n, const double* x, const int* pairs, const double* pair_x, double dhat_sq,
                   double kappa);

  private:
    double sum(hipStream_t s, const double* v, long n);
    DevBuf<double> buf_, part_;
};

}  // namespace sgipc
