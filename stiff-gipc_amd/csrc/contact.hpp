// Contact-pair generation (LBVH broad phase + distance-type narrow phase) and
// the barrier gradient/Hessian of every pair type.
#pragma once
#include "sgipc_internal.hpp"

namespace sgipc {

// Linear BVH over triangles (k = 3) or edges (k = 2): Morton-sorted leaves
// (Karras 2012 radix tree), AABBs refitted bottom-up.  Node numbering: internal
// nodes [0, n - 1), leaf of sorted position s at n - 1 + s; the root is node 0
// (a single leaf when n == 1).
struct Lbvh {
    int n = 0;
    DevBuf<double> box;       // 6 per node: lo xyz, hi xyz
    DevBuf<double> leaf_box;  // 6 per primitive (primitive order)
    DevBuf<double> scene;     // 6 + partial boxes of the reduction
    DevBuf<uint64_t> keys, skeys;
    DevBuf<int> idx, sidx;    // sidx[s] = primitive at sorted position s
    DevBuf<int> child;        // 2 per internal node
    DevBuf<int> parent;       // per node (-1 at the root)
    DevBuf<int> flag;         // per internal node, refit arrivals
    DevBuf<unsigned char> temp;
    // dir (optional): boxes swept to x - alpha dir
    void build(hipStream_t s, const double* x, const uint32_t* prims, int n_prims, int k, const double* dir = nullptr,
               double alpha = 0.0);
};

// Counts of the reference's _cpNum: [0] pairs, [1] unused, [2] point-point,
// [3] point-edge, [4] point-triangle + edge-edge.
struct ContactCounts {
    int c[5] = {0, 0, 0, 0, 0};
};

// Max of a device array of values >= 1 (and the sum of a parallel int array).
class StepReducer {
  public:
    double max_inv(hipStream_t s, const double* v, const int* c, long n, long* total);

  private:
    DevBuf<double> part_;
    DevBuf<int> ipart_;
};

class ContactDetector {
  public:
    // Largest t in (0, 1] such that x - t dir keeps every candidate pair of
    // the surface apart (additive CCD with eta = 1 - slackness); candidates
    // from the boxes swept to x - alpha dir.
    double ccd_step(hipStream_t s, const double* x, const double* dir, double alpha, const uint32_t* faces, int n_faces,
                    const uint32_t* edges, int n_edges, const uint32_t* surf, int n_surf, const int* body_id,
                    const int* btype, double dhat_sq, double slackness, long* n_candidates);
    // Pairs within squared distance dhat_sq; writes at most `capacity` pairs
    // and their MatIndex.  Order: point-face pairs by (surface slot, face),
    // then edge-edge pairs by (edge, edge).  Throws SGIPC_CAPACITY (counts
    // filled) when capacity is short.
    ContactCounts detect(hipStream_t s, const double* x, const double* rest, int n_verts, const uint32_t* faces,
                         int n_faces, const uint32_t* edges, int n_edges, const uint32_t* surf, int n_surf,
                         const int* body_id, const int* btype, double dhat_sq, int* pairs, int* mat_index,
                         int capacity);
    // pair count of the last detection, also when it threw SGIPC_CAPACITY
    int last_total = 0;

  private:
    Lbvh fbvh_, ebvh_;
    DevBuf<int> cnt_, off_, err_;
    DevBuf<uint64_t> rkey_, rkey_sorted_;
    DevBuf<int> rid_, rid_sorted_;
    DevBuf<int4> rpair_;
    DevBuf<int> rcls_;
    DevBuf<int> flags_, scan_;
    DevBuf<unsigned char> temp_;
    DevBuf<double> vals_;
    StepReducer red_;
};

// Ground step (largest t keeping surface vertices above normal . x = offset
// with the reference's slackness rule) and tet injectivity step.
double ground_step(hipStream_t s, StepReducer& red, DevBuf<double>& vals, const double* x, const uint32_t* surf,
                   int n_surf, const double* normal, double offset, const double* dir, double slackness);
double injective_step(hipStream_t s, StepReducer& red, DevBuf<double>& vals, const double* x, const uint32_t* tets,
                      int n_tets, const double* dir, double slackness, double error_rate);

// Barrier gradient (atomically added into grad when non-null) and the
// Hessian blocks of every pair type into the contact arena
// [PT + EE x 16 | PE x 9 | PP x 4] by MatIndex (n4, n3: PT+EE and PE counts);
// lower blocks are moved to (col, row) and zeroed.
void launch_barrier_pairs(hipStream_t s, int n_pairs, const double* x, const int* pairs, const int* mat_index, int n4,
                          int n3, double dhat_sq, double kappa, int hess_offset, double* grad, double* tvals,
                          int* trows, int* tcols, int* bad);

}  // namespace sgipc
