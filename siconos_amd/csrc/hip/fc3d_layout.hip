// fc3d_layout.hip — host-side construction of the colour-major sliced layout
// (see fc3d_problem.hip for its description).  Several problems can be
// appended one after the other; each occupies whole slices.
#include <algorithm>
#include <cstdlib>

#include "fc3d_layout.hpp"

namespace fc3d {

int validate_rows(int c0, int n, const int64_t* row_ptr, const int64_t* col_idx, RowInfo* info,
                  std::string* err) {
  char buf[256];
  for (int i = c0; i < c0 + n; i++) {
    if (row_ptr[i + 1] < row_ptr[i]) {
      snprintf(buf, sizeof buf, "row pointers not monotone at row %d", i);
      *err = buf;
      return NUMERICS_ERR_INVALID_ARGUMENT;
    }
    int64_t d = -1;
    for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
      if (col_idx[k] < c0 || col_idx[k] >= c0 + n) {
        snprintf(buf, sizeof buf, "column %lld out of range in row %d", (long long)col_idx[k], i);
        *err = buf;
        return NUMERICS_ERR_INVALID_ARGUMENT;
      }
      if (col_idx[k] == i) {
        if (d >= 0) {
          snprintf(buf, sizeof buf, "two diagonal blocks in row %d", i);
          *err = buf;
          return NUMERICS_ERR_INVALID_ARGUMENT;
        }
        d = k;
      }
    }
    if (d < 0) {
      snprintf(buf, sizeof buf, "row %d has no diagonal block", i);
      *err = buf;
      return NUMERICS_ERR_INVALID_ARGUMENT;
    }
    info->diag_block[(size_t)i] = d;
    info->dpos[(size_t)i] = (int)(d - row_ptr[i]);
    info->offlen[(size_t)i] = (int)(row_ptr[i + 1] - row_ptr[i] - 1);
  }
  return 0;
}

int default_sort_window() {
  const char* w = getenv("FC3D_HIP_SORT_WINDOW");
  return w ? std::max(1, atoi(w)) : 256;
}

void append_problem(HostLayout& L, int c0, int n, const int64_t* row_ptr, const int64_t* col_idx,
                    const double* blocks, const double* q, const double* mu, const int* colour,
                    const RowInfo& info, int window, int* ncolours_out,
                    std::vector<int64_t>* block_dst) {
  int ncol = 0;
  for (int i = c0; i < c0 + n; i++) ncol = std::max(ncol, colour[i] + 1);
  *ncolours_out = ncol;
  // colour-major; inside a colour index order (keeps the caller's locality
  // for neighbour gathers), rows sorted by decreasing length inside windows
  // of `window` contacts to limit padding
  std::vector<std::vector<int>> members((size_t)ncol);
  for (int i = c0; i < c0 + n; i++) members[(size_t)colour[i]].push_back(i);
  const int slice0 = L.nslices;
  L.colour_slice_ptr.push_back(slice0);
  for (int c = 0; c < ncol; c++) {
    auto& m = members[(size_t)c];
    for (size_t w0 = 0; w0 < m.size(); w0 += (size_t)window)
      std::stable_sort(m.begin() + (long)w0,
                       m.begin() + (long)std::min(m.size(), w0 + (size_t)window),
                       [&](int a, int b) { return info.offlen[(size_t)a] > info.offlen[(size_t)b]; });
    L.nslices += (int)((m.size() + 63) / 64);
    L.colour_slice_ptr.push_back(L.nslices);
  }
  L.nslots = L.nslices * 64;
  L.slot2orig.resize((size_t)L.nslots, -1);
  if (L.orig2slot.size() < (size_t)(c0 + n)) L.orig2slot.resize((size_t)(c0 + n), -1);
  for (int c = 0; c < ncol; c++) {
    const auto& m = members[(size_t)c];
    const int sb = L.colour_slice_ptr[L.colour_slice_ptr.size() - (size_t)ncol - 1 + (size_t)c];
    double b64 = 0, b32 = 0;
    for (size_t t = 0; t < m.size(); t++) {
      const int slot = sb * 64 + (int)t;
      L.slot2orig[(size_t)slot] = m[t];
      L.orig2slot[(size_t)m[t]] = slot;
      const double kbar = info.offlen[(size_t)m[t]] + 1.0;
      b64 += 76.0 * kbar + 108.0;
      b32 += 40.0 * kbar + 56.0;
    }
    L.colour_bytes64.push_back(b64);
    L.colour_bytes32.push_back(b32);
  }
  // slice geometry
  const int nsl = L.nslices - slice0;
  L.koff.resize((size_t)L.nslices);
  L.kmax.resize((size_t)L.nslices);
  for (int sl = slice0; sl < L.nslices; sl++) {
    int mx = 0;
    for (int l = 0; l < 64; l++) {
      const int o = L.slot2orig[(size_t)sl * 64 + (size_t)l];
      if (o >= 0) mx = std::max(mx, info.offlen[(size_t)o]);
    }
    L.koff[(size_t)sl] = (int)L.nplanes;
    L.kmax[(size_t)sl] = mx;
    L.nplanes += mx;
  }
  (void)nsl;
  L.len.resize((size_t)L.nslots, -1);
  L.dpos.resize((size_t)L.nslots, 0);
  L.col.resize((size_t)L.nplanes * 64, 0);
  L.vals.resize((size_t)L.nplanes * 9 * 64, 0.0);
  L.diag.resize((size_t)L.nslices * 9 * 64, 0.0);
  L.q.resize((size_t)L.nslices * 3 * 64, 0.0);
  L.mu.resize((size_t)L.nslots, 0.0);
  // padded plane entries keep a zero block and point at the lane's own slot
  for (int sl = slice0; sl < L.nslices; sl++)
    for (int k = 0; k < L.kmax[(size_t)sl]; k++)
      for (int l = 0; l < 64; l++)
        L.col[((size_t)L.koff[(size_t)sl] + (size_t)k) * 64 + (size_t)l] = sl * 64 + l;
  if (block_dst) block_dst->assign((size_t)(row_ptr[c0 + n] - row_ptr[c0]), 0);
  for (int slot = slice0 * 64; slot < L.nslots; slot++) {
    const int i = L.slot2orig[(size_t)slot];
    if (i < 0) continue;
    const int sl = slot / 64, l = slot % 64;
    L.len[(size_t)slot] = info.offlen[(size_t)i];
    L.dpos[(size_t)slot] = info.dpos[(size_t)i];
    int kk = 0;
    for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
      if (k == info.diag_block[(size_t)i]) continue;
      const size_t plane = (size_t)L.koff[(size_t)sl] + (size_t)kk;
      L.col[plane * 64 + (size_t)l] = L.orig2slot[(size_t)col_idx[k]];
      for (int e = 0; e < 9; e++)
        L.vals[(plane * 9 + (size_t)e) * 64 + (size_t)l] = blocks[9 * k + e];
      if (block_dst)
        (*block_dst)[(size_t)(k - row_ptr[c0])] = (int64_t)(plane * 9 * 64 + (size_t)l);
      kk++;
    }
    const int64_t d = info.diag_block[(size_t)i];
    if (block_dst)
      (*block_dst)[(size_t)(d - row_ptr[c0])] = -1 - (int64_t)((size_t)sl * 9 * 64 + (size_t)l);
    for (int e = 0; e < 9; e++)
      L.diag[((size_t)sl * 9 + (size_t)e) * 64 + (size_t)l] = blocks[9 * d + e];
    for (int e = 0; e < 3; e++) L.q[((size_t)sl * 3 + (size_t)e) * 64 + (size_t)l] = q[3 * i + e];
    L.mu[(size_t)slot] = mu[i];
  }
}

void build_dataflow(const HostLayout& L, const int* colour, DataflowHost* out) {
  const int ns = L.nslices;
  std::vector<std::vector<int>> deps((size_t)ns);
  for (int s = 0; s < ns; s++) {
    const int kmax = L.kmax[(size_t)s];
    for (int lane = 0; lane < 64; lane++) {
      const int slot = s * 64 + lane;
      const int i = L.slot2orig[(size_t)slot];
      if (i < 0) continue;
      const int len = L.len[(size_t)slot];
      for (int k = 0; k < len && k < kmax; k++) {
        const int js = L.col[((size_t)L.koff[(size_t)s] + (size_t)k) * 64 + (size_t)lane];
        const int j = L.slot2orig[(size_t)js];
        const int sj = js / 64;
        if (j < 0 || sj == s) continue;
        if (colour[j] < colour[i])
          deps[(size_t)s].push_back(sj);
        else if (colour[j] > colour[i])
          deps[(size_t)sj].push_back(s);
      }
    }
  }
  out->dep_ptr.assign((size_t)ns + 1, 0);
  out->dep_idx.clear();
  for (int s = 0; s < ns; s++) {
    auto& d = deps[(size_t)s];
    std::sort(d.begin(), d.end());
    d.erase(std::unique(d.begin(), d.end()), d.end());
    out->dep_idx.insert(out->dep_idx.end(), d.begin(), d.end());
    out->dep_ptr[(size_t)s + 1] = (int)out->dep_idx.size();
  }
  // per-XCD queues, colour-major inside each queue
  std::vector<std::vector<int>> q(8);
  const int ncol = (int)L.colour_slice_ptr.size() - 1;
  for (int c = 0; c < ncol; c++) {
    const int sb = L.colour_slice_ptr[(size_t)c], se = L.colour_slice_ptr[(size_t)c + 1];
    const int n = se - sb;
    for (int x = 0; x < 8; x++)
      for (int t = sb + (int)((int64_t)n * x / 8); t < sb + (int)((int64_t)n * (x + 1) / 8); t++)
        q[(size_t)x].push_back(t);
  }
  out->qlist.clear();
  for (int x = 0; x < 8; x++) {
    out->qoff[x] = (int)out->qlist.size();
    out->qlist.insert(out->qlist.end(), q[(size_t)x].begin(), q[(size_t)x].end());
  }
  out->qoff[8] = (int)out->qlist.size();
}

}  // namespace fc3d
