// Vertex <-> MAS slot remapping of a CEMAS partition.  With a partition the
// system is solved in slot space: vertex i becomes row slot_of[i], the slots
// a part does not fill are empty rows (no blocks, zero right-hand side), so
// the MAS sees the reference's slot layout and the Krylov iterates on the
// filled slots are those of the vertex-space system.
#pragma once
#include "sgipc_internal.hpp"

namespace sgipc {

// In place over n upper triplets: (r, c, H) -> (s_r, s_c, H) when s_r <= s_c,
// else (s_c, s_r, H^T), with s = slot_of[] (an entry outside [0, n_verts)
// gets row -1, which the converter's index check reports).
void launch_slot_triplets(hipStream_t s, long n, int n_verts, const int* slot_of, double* vals, int* rows, int* cols);
// Contact pairs in the reference's int4 encodings, vertex ids renamed to
// slots (sign encodings and the non-vertex entries kept).
void launch_slot_pairs(hipStream_t s, int n, const int* slot_of, const int* pairs, int* out);
// out (3 n_slots) = v (3 n) placed at the slots, zero elsewhere.
void launch_slot_scatter(hipStream_t s, int n, int n_slots, const int* slot_of, const double* v, double* out);
// v (3 n) = in (3 n_slots) read at the slots.
void launch_slot_gather(hipStream_t s, int n, const int* slot_of, const double* in, double* v);

}  // namespace sgipc
