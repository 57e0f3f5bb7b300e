// fc3d_layout.hpp — host-side builder of the colour-major sliced layout.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "fc3d_numerics.h"

namespace fc3d {

// Per-row facts of the caller's CSR (indexed by global contact).
struct RowInfo {
  std::vector<int> offlen;          // off-diagonal blocks of the row
  std::vector<int> dpos;            // position of the diagonal block in the row
  std::vector<int64_t> diag_block;  // index of the diagonal block
  explicit RowInfo(size_t n) : offlen(n), dpos(n), diag_block(n) {}
};

// Host image of the device layout; problems are appended slice-aligned.
struct HostLayout {
  int nslices = 0, nslots = 0;
  int64_t nplanes = 0;
  std::vector<int> len, dpos, koff, kmax, col;
  std::vector<double> vals, diag, q, mu;
  std::vector<int> slot2orig, orig2slot;
  std::vector<int> colour_slice_ptr;  // per appended problem: ncolours + 1 entries
  std::vector<double> colour_bytes64, colour_bytes32;  // per colour, algorithmic bytes
};

// Checks rows [c0, c0 + n): monotone pointers, columns inside the range, one
// diagonal block per row.
int validate_rows(int c0, int n, const int64_t* row_ptr, const int64_t* col_idx, RowInfo* info,
                  std::string* err);
int default_sort_window();
// block_dst (optional, one entry per CSR block of the problem, indexed from
// row_ptr[c0]): where block k's first value lands, as (plane * 9 * 64 + lane)
// in `vals` for an off-diagonal block or as -1 - (slice * 9 * 64 + lane) in
// `diag` for the diagonal one; value e then sits 64 entries further per e.
void append_problem(HostLayout& L, int c0, int n, const int64_t* row_ptr, const int64_t* col_idx,
                    const double* blocks, const double* q, const double* mu, const int* colour,
                    const RowInfo& info, int window, int* ncolours_out,
                    std::vector<int64_t>* block_dst = nullptr);

// Dependencies of the single-launch (dataflow) sweep.  A slice may start once
// every slice holding a neighbour of lower colour has finished the current
// sweep; neighbours of higher colour wait for it in turn (edges are taken
// from both directions of the sparsity, so a non-symmetric pattern is safe).
// Slices are dealt from one queue per XCD; each queue lists, colour after
// colour, a contiguous eighth of the colour's slices.
struct DataflowHost {
  std::vector<int> dep_ptr;  // nslices + 1
  std::vector<int> dep_idx;
  std::vector<int> qlist;    // nslices: the queues one after the other
  int qoff[9];               // queue q = qlist[qoff[q] .. qoff[q + 1])
};
void build_dataflow(const HostLayout& L, const int* colour, DataflowHost* out);

}  // namespace fc3d
