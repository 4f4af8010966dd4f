// Multilevel additive Schwarz preconditioner (host object).
#pragma once
#include "assembly.hpp"
#include "mpnn.hpp"
#include "pcg.hpp"
#include "sgipc_internal.hpp"
#include "spmv_full.hpp"

namespace sgipc {

constexpr int kMasMaxLevels = 6;
// layout of the device metadata array
constexpr int kMetaCnt = 0;   // counts c_0..c_L
constexpr int kMetaOff = 8;   // offsets off_0..off_L
constexpr int kMetaSize = 32;

class Mas {
  public:
    Mas(Context* ctx, int bank);
    ~Mas();
    Mas(const Mas&) = delete;
    Mas& operator=(const Mas&) = delete;
    void init_neighbor(int n, int coff, int total_nbr, const uint32_t* start, const uint32_t* num,
                       const uint32_t* list);
    void init_matrix();
    void set_bcoo(const double* vals, const int* rows, const int* cols, int nnz, int offset,
                  const int* pairs, int n_pairs, int precision);
    // The hierarchy of the next set_bcoo (it depends on the connectivity and
    // the contact pairs only), built ahead on stream s, e.g. while the caller
    // assembles the matrix; returns after its work on s is complete.  The
    // next set_bcoo with the same pairs pointer and count uses it.
    void prebuild_hierarchy(hipStream_t s, const int* pairs, int n_pairs);
    // z = M^-1 r, rz partials; see Preconditioner.
    template <typename TV>
    int apply(hipStream_t s, const TV* r, TV* z, double* rz_part, const PcgStopDesc& sd);
    // x/r update fused into the restriction; needs L > 2 (the stop rule is
    // evaluated by the first coarse restriction).  Returns -1 otherwise.
    template <typename TV>
    int apply_update_xr(hipStream_t s, const FusedXR<TV>& fx, TV* z, double* rz_part, int* n_rr);
    // rr partials apply_update_xr writes
    int rr_partials() const { return (h_off[1] + 255) / 256; }
    int levels(int* offsets, int* counts) const;
    bool level2_exact() const { return dense_m_ > 0; }
    // 0: additive levels (MASPreconditioner::preconditioning); 1: a symmetric
    // V(1,1) cycle on level 1 (level-1 domain solves as smoother, the levels
    // >= 2 as its coarse correction) with the level-0 domain solves additive;
    // 2: mode 1's level-1 cycle as the coarse solve of a symmetric two-grid
    // cycle on level 0 (level-0 domain solves weighted by omega0 before and
    // after it; needs the fine operator).  Takes effect at the next set_bcoo.
    int coarse_mode = 0;
    double omega0 = 0.4;
    // coarse mode 2 with the float apply: the level-0 domain solves read the
    // inverse as FP16 scaled to a unit diagonal (C = D^-1 Minv D^-1, D^2 its
    // diagonal) and the per-slot scales, half the bytes of the float blocks.
    // Takes effect at the next set_bcoo.
    bool half_level0 = true;
    // the fine operator of coarse mode 2 (the PCG matrix, with its FP32 copy
    // when the apply is FP32); set_bcoo builds its own unless told not to
    const FullMatrix* fine_ = nullptr;
    bool build_own_fine = true;
    bool vcycle_active(bool f32) const
    {
        return coarse_mode == 2 && cycle_active() && ap_nnz_ > 0 && fine_ != nullptr && fine_->n_rows == n &&
               (!f32 || fine_->has_f32);
    }
    // the coarse path the last apply took: 2 the two-grid cycle, 1 the
    // level-1 cycle, 0 the additive levels (-1 before the first apply).  A
    // configured mode can fall back silently (no fine operator, one level-1
    // domain, no FP32 copy), so solves report this, not coarse_mode.
    int last_path() const { return last_path_; }
    bool keep_graph = false;  // build the level-1 graph also without the cycle
    // flexible PCG: z.Ap partials written by the prolongation when set
    const void* flex_Ap_ = nullptr;
    double* flex_zap_ = nullptr;
    // learned coarse correction (nullptr: off); needs the level-1 graph
    Mpnn* net_ = nullptr;
    bool learned_active() const
    {
        return net_ != nullptr && graph_.n == h_cnt[1] && h_cnt[1] > 0 && L > 1;
    }
    // training sample of residual r: features X (3 c_1 rows x 3) and target q
    // (3 c_1), normalised by the RMS of the first feature; ztmp: 3n scratch
    void learned_sample(hipStream_t s, const double* r, double* ztmp, float* X, float* q);
    const CoarseGraph& level1_graph() const { return graph_; }
    void tables(int* coarse_table, int* going_next) const;
    void inverse_blocks(double* out) const;
    // the dense level-2 inverse of the level-1 cycle as m x m row-major into
    // out (device; nullptr: size only); returns m (0: no exact level-2 solve)
    int level2_inverse(double* out) const;

    Context* ctx;
    int B;
    int n = 0, L = 0, coff = 0, precision = 0;
    int h_cnt[kMasMaxLevels + 2] = {0};
    int h_off[kMasMaxLevels + 2] = {0};
    int total = 0;
    bool built = false;
    double ms_build = 0.0;

  private:
    void build_hierarchy(hipStream_t s, const int* pairs, int n_pairs);
    template <int BB>
    void build_hierarchy_t(hipStream_t s, const int* pairs, int n_pairs);
    bool hier_ready_ = false;  // prebuilt for (hier_pairs_, hier_npairs_)
    const int* hier_pairs_ = nullptr;
    int hier_npairs_ = 0;
    template <int BB>
    void assemble_invert_t(const double* vals, const int* rows, const int* cols, int nnz, int offset,
                           bool want_ap);
    template <int BB, typename TV, typename TS>
    int apply_t(hipStream_t s, const TV* r, TV* z, double* rz_part, const PcgStopDesc& sd);
    template <int BB, typename TV, typename TS>
    int apply_xr_t(hipStream_t s, const FusedXR<TV>& fx, TV* z, double* rz_part, int* n_rr);
    // levels >= 1 (additive, or the level-1 cycle) and the prolongation
    template <int BB, typename TV, typename TS>
    int finish_t(hipStream_t s, TS* R, TS* Z, const TS* minv, const int* done, const PcgStopDesc& sd2, const TV* r,
                 TV* z, double* rz_part);
    // coarse mode 2 from the level-0 slots of r on
    template <int BB, typename TV, typename TS>
    int finish_v_t(hipStream_t s, TS* R, TS* Z, const TS* minv, const int* done, const PcgStopDesc& sd2,
                   const TV* r, TV* z, double* rz_part);
    bool want_graph() const { return coarse_mode >= 1 || keep_graph; }
    int last_path_ = -1;
    void note_path(bool f32) { last_path_ = vcycle_active(f32) ? 2 : (cycle_active() ? 1 : 0); }
    // the cycle needs level 1 to span several domains (one domain is solved
    // exactly already) and a coarser level
    bool cycle_active() const
    {
        return coarse_mode >= 1 && L > 2 && graph_.n == h_cnt[1] && h_cnt[1] > B;
    }
    // A P for the post-smoothing residual of coarse mode 2 (CSR over the fine
    // rows, level-1 columns, FP32 3x3 blocks)
    void build_ap(hipStream_t s, const double* vals, const int* rows, const int* cols, int nnz);
    DevBuf<int> ap_ptr_, ap_col_, ap_seg_;
    // with a row offset (the MAS covering the rows past the body rows): the
    // BCOO's suffix of rows >= offset, its rows and columns shifted to the
    // MAS's own numbering (A P and the own fine operator read it)
    DevBuf<int> sub_rows_, sub_cols_;
    long sub_e0_ = 0;
    int sub_nnz_ = 0;
    DevBuf<float> ap_blk_;
    DevBuf<double4> c4_;  // level-1 correction, 4-padded (float4 view in FP32 applies)
    int ap_nnz_ = 0;
    DevBuf<uint64_t> apk_, apk_sorted_;
    DevBuf<int> api_, api_sorted_, apf_, apinc_;
    DevBuf<unsigned char> ap_temp_;
    FullMatrix own_fine_;

    void build_dense_level2(hipStream_t s, const double* v, const int* r, const int* c, int nb);
    void finish_dense_level2(hipStream_t s);  // reads the inverse's pivot check
    static constexpr long kDenseMax = 4608;  // largest dense level-2 system (3 c_2)
    DevBuf<uint16_t> minv16_;    // level-0 inverse, FP16 scaled form (half_level0)
    DevBuf<float> msc_;          // its per-slot scales
    bool half_ready_ = false;
    DevBuf<double> dense_inv_;   // inverse of the level-2 operator (FP64, m x m in ld x ld)
    DevBuf<double> dense_piv_;   // blocked Gauss-Jordan: inverse of the pivot tile
    DevBuf<double> dense_col_;   // blocked Gauss-Jordan: the pivot column's tiles
    DevBuf<int> dense_info_;     // nonzero: a pivot was not positive
    int dense_m_ = 0;
    int dense_ld_ = 0;
    int dense_m_pending_ = 0;
    bool dense_pending_ = false;

    CoarseGraph graph_;          // level-1 operator (coarse cycle, learned level)
    DevBuf<double> g1d_, c2d_;   // level-1 cycle vectors (FP64 apply)
    DevBuf<float> g1f_, c2f_;    // level-1 cycle vectors (FP32 apply)
    DevBuf<float> lX_, lchat_, lcorr_;  // learned correction: features, output, scaled output

    DevBuf<uint32_t> nbr_start_, nbr_num_, nbr_list_;
    DevBuf<int> map_;        // [L][n], row 0 unused (identity)
    DevBuf<int> meta_;
    DevBuf<int> going_next_;
    DevBuf<uint32_t> mask_;
    DevBuf<int> comp_rank_, dom_count_, dom_prefix_, next_id_;
    DevBuf<int> row_ptr_;
    DevBuf<double> mgal_;     // coarse-level Galerkin blocks
    DevBuf<double> minv64_;
    DevBuf<float> minv32_;
    DevBuf<double> R64_, Z64_;
    DevBuf<float> R32_, Z32_;
    // coarse operators (ping-pong) and mapped triplets of the coarsening
    DevBuf<double> coarse_v_[2], map_v_;
    DevBuf<int> coarse_r_[2], coarse_c_[2], map_r_, map_c_;
    Converter conv_;
    SideStream side_;  // A P and the level-2 elimination, beside the coarse levels and the domain inverses
};

// Adapter giving the PCG driver a preconditioner over a Mas.
template <typename TV>
struct MasPrecond : Preconditioner<TV> {
    Mas* mas;
    explicit MasPrecond(Mas* m) : mas(m) {}
    int apply(hipStream_t s, const TV* r, TV* z, double* rz_part, const PcgStopDesc& sd) override
    {
        return mas->template apply<TV>(s, r, z, rz_part, sd);
    }
    int apply_update_xr(hipStream_t s, const FusedXR<TV>& fx, TV* z, double* rz_part, int* n_rr) override
    {
        return mas->template apply_update_xr<TV>(s, fx, z, rz_part, n_rr);
    }
    int max_rr_partials() const override { return mas->rr_partials(); }
    bool set_flex(const TV* Ap, double* zap) override
    {
        mas->flex_Ap_ = Ap;
        mas->flex_zap_ = zap;
        return true;
    }
};

int mas_num_levels(int n, int bank);

}  // namespace sgipc
