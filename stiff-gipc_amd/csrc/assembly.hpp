// Hessian assembly launchers and the triplet -> BCOO converter.
#pragma once
#include "sgipc_internal.hpp"

namespace sgipc {

void launch_tet_snh(hipStream_t s, int n_tets, const double* x, const uint32_t* tets, const double* dm_inv,
                    const double* vol, const double* mu, const double* lam, double dt, int hess_offset,
                    double* grad, double* tvals, int* trows, int* tcols);
void launch_barrier_pt(hipStream_t s, int n_pairs, const double* x, const int* pairs, const double* pair_x,
                       double dhat_sq, double kappa, int hess_offset, double* grad, double* tvals, int* trows,
                       int* tcols);
void launch_tri_bw(hipStream_t s, int n_tris, const double* x, const uint32_t* tris, const double* dm_inv2,
                   const double* area, double k_stretch, double k_shear, double strain_rate, double dt,
                   int hess_offset, double* grad, double* tvals, int* trows, int* tcols);
void launch_quad_bending_q(hipStream_t s, int n_edges, const double* rest, const uint32_t* edges,
                           const uint32_t* adj, double* Q16);
void launch_quad_bending(hipStream_t s, int n_edges, const double* x, const double* rest, const uint32_t* edges,
                         const uint32_t* adj, const double* Q16, double k_bend, double dt, int hess_offset,
                         double* grad, double* tvals, int* trows, int* tcols);
void launch_mass_diag(hipStream_t s, int n, const double* mass, int hess_offset, double* tvals, int* trows,
                      int* tcols);

// The device scratch one conversion of n triplets works in: sort keys and
// their sorted copy, triplet indices and their sorted copy, head flags and
// their inclusive scan, block starts (n + 1 entries), 4 ints, and the radix
// sort / scan temporary storage.
struct ConvertScratch {
    uint64_t* keys = nullptr;
    uint64_t* skeys = nullptr;
    int* idx = nullptr;
    int* sidx = nullptr;
    int* flag = nullptr;
    int* pos = nullptr;
    int* head = nullptr;
    int* ints = nullptr;
    void* temp = nullptr;
    size_t temp_bytes = 0;
};
// Temporary storage (bytes) the radix sort and scan of n triplets need.
size_t convert_temp_bytes(hipStream_t s, long n);
// Largest row/col index of n triplets (ints[0] is the scratch word).
int convert_max_index(hipStream_t s, const int* rows, const int* cols, long n, int* ints);
// Sort + reduce-by-key into the caller's scratch; writes at most `capacity`
// blocks and stores the unique count in *unique (also when it throws
// SGIPC_CAPACITY).  Allocates nothing.
int convert_into(hipStream_t s, const ConvertScratch& w, const double* tvals, const int* trows, const int* tcols,
                 long n, int index_bound, double* ovals, int* orows, int* ocols, int capacity, int* unique);

// Sort + reduce-by-key of 3x3 triplets into unique upper BCOO.  Buffers are
// kept between calls.
class Converter {
  public:
    // Returns the unique block count; throws SGIPC_CAPACITY / SGIPC_BAD_INDEX.
    // index_bound: every row/col is < index_bound.
    int convert(hipStream_t s, const double* tvals, const int* trows, const int* tcols, long n, int index_bound,
                double* ovals, int* orows, int* ocols, int capacity);
    int max_index(hipStream_t s, const int* rows, const int* cols, long n);
    static int index_bits(int bound);
    // unique block count of the last convert, also when it threw
    // SGIPC_CAPACITY (the capacity the caller must provide)
    int last_unique = 0;

  private:
    DevBuf<uint64_t> keys_, skeys_;
    DevBuf<int> idx_, sidx_, flag_, pos_, head_, scratch_int_;
    DevBuf<unsigned char> temp_;
};

}  // namespace sgipc
