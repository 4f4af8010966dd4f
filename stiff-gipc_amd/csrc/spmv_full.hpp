// Symmetric matrix expanded to both triangles and laid out as SELL-64-sigma for
// the PCG SpMV.
#pragma once
#include "sgipc_internal.hpp"

namespace sgipc {

class FullMatrix {
  public:
    // Expand a unique upper BCOO (rows sorted, row <= col, indices < n) into
    // both triangles and pack it as SELL: rows stably sorted by length, 64 rows
    // per slice, each slice as wide as its longest row, block values stored
    // component-major per slice.  want_f32 also keeps FP32 values.
    void build(hipStream_t s, const double* vals, const int* rows, const int* cols, int nnz_up, int n,
               bool want_f32);
    // y = a*A*x + b*y; dot_part (optional) receives per-workgroup x^T(A x)
    // partials; returns the partial count.  Returns at once when *done != 0.
    template <typename TV>
    int apply(hipStream_t s, TV a, const TV* x, TV b, TV* y, double* dot_part, const int* done) const;
    int partials() const;

    int n_rows = 0;
    int nnz = 0;         // blocks of the expanded matrix
    long n_slots = 0;    // SELL slots (64 blocks each, padding included)
    int n_slices = 0;
    bool has_f32 = false;
    DevBuf<double> sell_vals;
    DevBuf<float> sell_vals32;
    DevBuf<int> sell_cols, slice_ptr, slice_w, perm;

  private:
    DevBuf<double> fvals_;
    DevBuf<int> frows_, fcols_, row_ptr_;
    DevBuf<uint64_t> keys_, skeys_;
    DevBuf<int> idx_, sidx_, up_ptr_, col_ptr_, low_cnt_, full_len_, len_key_, len_key_sorted_, iota_;
    DevBuf<unsigned char> temp_;
};

}  // namespace sgipc
