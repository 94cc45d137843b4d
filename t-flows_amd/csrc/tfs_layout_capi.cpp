// tfs_layout_capi.cpp -- the solver's host-side layout builders behind a C
// ABI in the CPU library (libtfs_mesh.so), so that the permutations and level
// structures the device code relies on can be checked without a GPU.
#include <cstddef>
#include <vector>

#include "tfs_layout.hpp"
#include "tfs_mesh.h"

extern "C" int tfs_layout_block_colour(int n, const int *row, const int *col, int block,
                                       int *perm, int *colour) {
  if (n < 0 || block < 1 || (n > 0 && (!row || !col || !perm || !colour))) return -1;
  std::vector<int> p, c;
  const int ncol = tfs::colour_order(n, row, col, p, &c, block);
  for (int i = 0; i < n; ++i) {
    perm[i] = p[i];
    colour[i] = c[i];
  }
  return ncol;
}

extern "C" int tfs_layout_colour(int n, const int *row, const int *col, int *perm, int *colour) {
  if (n < 0 || (n > 0 && (!row || !col || !perm || !colour))) return -1;
  std::vector<int> p, c;
  const int ncol = tfs::colour_order(n, row, col, p, &c);
  for (int i = 0; i < n; ++i) {
    perm[i] = p[i];
    colour[i] = c[i];
  }
  return ncol;
}

extern "C" int tfs_layout_summary(int n, const int *row, const int *col, const int *dia,
                                  int *s_orig, int *level_lower, int *level_upper,
                                  int *bandwidth) {
  if (n < 1 || !row || !col || !dia || !s_orig || !level_lower || !level_upper || !bandwidth)
    return -1;
  tfs::Layout L;
  tfs::build_layout(n, row, col, dia, 64, 512, 2048, L);
  for (int i = 0; i < n; ++i) s_orig[i] = L.s_orig[i];
  for (const tfs::SweepLayout *S : {&L.L, &L.U}) {
    int *lev = S == &L.L ? level_lower : level_upper;
    for (int l = 0; l < S->nlev; ++l)
      for (int q = S->lev_ptr[l]; q < S->lev_ptr[l + 1]; ++q) lev[S->orig[q]] = l;
  }
  *bandwidth = L.bandwidth;
  return 0;
}

extern "C" int tfs_layout_block_colour_bfs(int n, const int *row, const int *col, int block,
                                           int *perm, int *colour, int *block_of) {
  if (n < 0 || block < 1 || (n > 0 && (!row || !col || !perm || !colour || !block_of)))
    return -1;
  std::vector<int> p, c, b;
  const int ncol = tfs::block_colour_order(n, row, col, block, p, &c, &b);
  for (int i = 0; i < n; ++i) {
    perm[i] = p[i];
    colour[i] = c[i];
    block_of[i] = b[i];
  }
  return ncol;
}

extern "C" int tfs_layout_brick_colour(int n, const int *row, const int *col, const int *dia,
                                       int edge, int cap, int *perm, int *colour,
                                       int *block_of) {
  if (n < 0 || edge < 1 || cap < 1 ||
      (n > 0 && (!row || !col || !dia || !perm || !colour || !block_of)))
    return -1;
  std::vector<int> p, bstart, bcolour;
  const int ncol = tfs::brick_colour_order(n, row, col, dia, edge, cap, p, bstart, bcolour);
  for (std::size_t b = 0; b + 1 < bstart.size(); ++b)
    for (int k = bstart[b]; k < bstart[b + 1]; ++k) {
      perm[k] = p[(size_t)k];
      colour[p[(size_t)k]] = bcolour[b];
      block_of[p[(size_t)k]] = (int)b;
    }
  return ncol;
}
