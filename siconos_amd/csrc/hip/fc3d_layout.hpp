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
void append_problem(HostLayout& L, int c0, int n, const int64_t* row_ptr, const int64_t* col_idx,
                    const double* blocks, const double* q, const double* mu, const int* colour,
                    const RowInfo& info, int window, int* ncolours_out);

}  // namespace fc3d
