// Symmetric matrix laid out for the PCG SpMV: row tiles owned by one
// workgroup, in-tile upper blocks stored once (transposed half summed in LDS),
// cross-tile blocks stored in both rows, SELL-64 packing inside each tile.
#pragma once
#include "sgipc_internal.hpp"

namespace sgipc {

// Level-1 sums of a MAS level-0 residual held in the same slot array (coarse
// mode 2): for every row u < n_rows, y[npad + map1[u]] += y[u] (3 components,
// summed per level-1 node in a fixed order), and the level-1 padding slots
// [npad + c1, npad + c1 + pad) set to zero.
struct Level1Sums {
    const int* map1 = nullptr;
    int npad = 0;
    int c1 = 0;
    int pad = 0;
};

class FullMatrix {
  public:
    // Lay out a unique upper BCOO (rows sorted, row <= col, indices < n).
    // want_f32 also keeps FP32 values.
    void build(hipStream_t s, const double* vals, const int* rows, const int* cols, int nnz_up, int n,
               bool want_f32);
    // y = a*A*x + b*y; dot_part (optional) receives one x^T(A x) partial per
    // tile; returns the partial count (0 without dot_part).  Returns at once
    // when *done != 0.
    template <typename TV>
    int apply(hipStream_t s, TV a, const TV* x, TV b, TV* y, double* dot_part, const int* done) const;
    // y = a*A*x + b*y followed by the level-1 sums of y's rows (l1), in one
    // launch of the halo form (the epilogue holds whole 512-row tiles, so
    // whole level-0 domains).  Returns false, having launched nothing, when
    // the matrix is not in the halo form.
    template <typename TV>
    bool apply_level1_sums(hipStream_t s, TV a, const TV* x, TV b, TV* y, const int* done, const Level1Sums& l1) const;
    // FP64 vectors against the FP32 copy of the values (half the matrix
    // bytes), FP64 accumulation; needs want_f32 at build.
    int apply_f32_values(hipStream_t s, double a, const double* x, double b, double* y, double* dot_part,
                         const int* done) const;
    // PCG iteration with the p update fused in: p_new = z + beta p_old
    // (beta = rz_new / rz_old, summed from n_rz partials) is formed while it
    // is gathered, written for every owned row, and Ap = A p_new.
    template <typename TV>
    int apply_update_p(hipStream_t s, bool f32_values, const TV* z, const TV* p_old, TV* p_new, TV* Ap,
                       double* dot_part, const int* done, const double* rz_new, const double* rz_old,
                       int n_rz) const;
    // partial count of apply() with dot_part
    int partials() const;
    // blocks stored in the layout (upper blocks + cross-tile transposed
    // copies, padding excluded); synchronises the stream
    long stored_blocks(hipStream_t s) const;

    int n_rows = 0;
    int nnz = 0;         // unique upper blocks of the source matrix
    bool use_halo = false;  // words are local indices (halo kernel); else global columns
    int halo_max = 0;       // largest halo of a tile (vertices)
    int tl_max = 0;         // in-tile transposed slots the fullest tile uses
    long n_slots = 0;    // SELL slots (64 entries each, padding included)
    int n_slices = 0;
    int n_tiles = 0;
    bool has_f32 = false;
    // rows with too many blocks for the tiles: their full rows as CSR
    // (column order, FP64 and, with want_f32, FP32 values), multiplied by a
    // kernel of their own after the tiled one; partials() counts one x^T(Ax)
    // partial per long row after the tiles'
    int n_long = 0;
    DevBuf<int> long_rows, long_ptr, long_col;
    DevBuf<double> long_vals;
    DevBuf<float> long_vals32;
    DevBuf<double> sell_vals;
    DevBuf<float> sell_vals32;
    DevBuf<uint32_t> sell_cols, tinfo;
    DevBuf<int> slice_ptr, slice_w, perm, plen;
    DevBuf<int> halo_ptr, halo_list, order;  // per-tile halo lists; launch order of the tiles
    // when both are set, every SpMV launch records them at its own start and
    // end (hipExtLaunchKernelGGL): the kernel's duration without the
    // dispatch overhead that separate event records add
    hipEvent_t time_start = nullptr, time_stop = nullptr;

    // dynamic LDS of the halo kernel with LDS entries of tv_size bytes and
    // `threads` threads per workgroup
    size_t halo_dyn_bytes(size_t tv_size, int threads) const;
    // largest static LDS among the FP64-vector halo kernels (queried once)
    static size_t halo_static_lds();

  private:
    void build_halo(hipStream_t s, int tbits);
    template <typename TM, typename TV>
    void launch_long(hipStream_t s, TV a, const TV* x, TV* y, double* dot_part, const int* done) const;
    template <typename TM, typename TV>
    int launch(hipStream_t s, const TM* vals, TV a, const TV* x, TV b, TV* y, double* dot_part, const int* done,
               const TV* z, TV* p_new, const double* rzn, const double* rzo, int n_rz,
               const Level1Sums* l1 = nullptr) const;

    DevBuf<uint64_t> keys_, skeys_, rkey_, rkey_sorted_;
    DevBuf<int> idx_, sidx_, cand_, cscan_, up_ptr_, col_ptr_, low_cnt_, len_, iota_, sorted_;
    DevBuf<uint32_t> up_word_;
    DevBuf<unsigned char> temp_;
    DevBuf<uint64_t> hkeys_, hkeys_sorted_;
    DevBuf<int> hcnt_, hoff_, hflag_, hinc_, stats_;
    DevBuf<int> lflag_, llen_, lrank_, loff_;
};

}  // namespace sgipc
