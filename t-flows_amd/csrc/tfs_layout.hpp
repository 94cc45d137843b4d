// tfs_layout.hpp -- device data layouts of one CRS matrix (host-side build).
//
// The solver keeps the matrix in three row orders, each chosen for the
// memory access pattern of the kernels that use it.  None changes any
// arithmetic: rows keep their entries in the caller's column order, so every
// row sum is formed exactly as the reference forms it.
//
//   S  (SpMV and the Krylov vectors): reverse Cuthill-McKee order of the
//      pattern, so that the x gathers of a wavefront's rows fall in a narrow
//      band (L2-resident) even when the caller numbers cells randomly.
//   L  (factor and forward sweep): rows grouped by level of the lower
//      dependency DAG of the CALLER's numbering (which defines the
//      incomplete factor), ties in S order; only the strictly-lower entries
//      are stored, contiguously, level after level.
//   U  (backward sweep): the L order backwards, strictly-upper entries.  A
//      row's upper dependencies sit at deeper lower levels, so they come
//      first; the backward sweep reads the forward sweep's output in runs.
#pragma once

#include <vector>

namespace tfs {

struct Segment {
  int l0, l1;  // levels [l0, l1)
  int rows;
  bool small;  // a run of small levels walked by one workgroup
};

struct SweepLayout {
  int nlev = 0;
  std::vector<int> lev_ptr;   // nlev+1 position starts of the levels
  std::vector<int> ptr;       // n+1 offsets of each position's triangle entries
  std::vector<int> col;       // neighbour positions in this layout
  std::vector<int> src;       // index of the entry in the caller's val[]
  std::vector<int> diag_src;  // index of the row's diagonal in val[]
  std::vector<int> orig;      // caller's row number of each position
  std::vector<int> of_orig;   // position of each caller row
  std::vector<Segment> segs;
};

struct Layout {
  int n = 0;
  long long nnz = 0;
  // S layout
  std::vector<int> s_orig, s_of_orig;   // S position <-> caller row
  std::vector<int> s_row, s_col, s_src, s_diag_src;
  std::vector<int> s_rblk;              // SpMV row blocks over S rows
  SweepLayout L, U;
  std::vector<int> l_to_s;  // S position of the row at L position p
  std::vector<int> u_to_l;  // L position of the row at U position q
  std::vector<int> s_to_u;  // U position of the row at S position k
  int bandwidth = 0;        // max |S(i) - S(j)| over entries (diagnostic)
};

// Greedy colouring of the (symmetrised) pattern, vertices visited in index
// order, each taking the smallest colour none of its neighbours has.
// `perm` lists the vertices by colour, index order within a colour (new
// position -> vertex); returns the number of colours.  Rows of one colour
// are independent, so an incomplete factor in this order has one level per
// colour.  With block > 1, runs of `block` consecutive vertices are coloured
// as units (block multi-colouring): a block keeps its internal couplings, in
// order, and an incomplete factor has at most `block` levels per colour.
int colour_order(int n, const int *row, const int *col, std::vector<int> &perm,
                 std::vector<int> *colour = nullptr, int block = 1);

// Algebraic block multi-colour order (Iwashita, Nakashima & Takahashi,
// "Algebraic Block Multi-Color Ordering Method for Parallel Multi-Threaded
// Sparse Triangular Solver in ICCG Method", IPDPS 2012): blocks of at most
// `block` rows grown breadth-first over the symmetrised pattern (seeds in
// index order, a block's rows in the order the search reached them, neighbours
// in ascending index), the block graph coloured greedily in block order, and
// the rows listed by colour, then block, then discovery order.  An incomplete
// factor in this order couples a row only to earlier rows of its own block and
// to rows of blocks of earlier colours: a sweep is one pass per colour whose
// blocks are independent, each block walked in its own (short) level order.
// `perm` maps new position -> vertex; `colour` and `block_of` are per vertex
// (block numbers in growth order).  Returns the number of colours.
int block_colour_order(int n, const int *row, const int *col, int block, std::vector<int> &perm,
                       std::vector<int> *colour = nullptr, std::vector<int> *block_of = nullptr);

// row/col/dia are 0-based, validated (diagonal present, columns ascending).
// identity_s keeps the caller's order as the S order (for a matrix whose rows
// are already in the order its kernels want, e.g. a block multi-colour copy).
// Block multi-colour order for the FP32 inner solves, from the caller's
// numbering: rows grouped by (phi_0 / edge, phi_1 / edge, phi_2 / edge) of the
// monotone potentials of the lower triangle's dependency graph -- phi_c(i) is
// the longest path to i counting the dependencies j -> i (j < i) of class c,
// the classes being the binary orders of the index step i - j merged into
// three groups of about equal entry counts; on a lexicographically numbered
// grid (steps 1, nx, nx ny) they are the cell's coordinates, so the groups are
// edge^3 bricks -- groups larger than `cap` rows split in caller order, and the
// blocks coloured greedily (in order of their first row) so that no two
// adjacent blocks share a colour.  perm lists the rows by (colour, block,
// caller order); bstart (nblocks + 1) gives each block's first position in
// that order and bcolour its colour.  Returns the number of colours.
// row/col/dia: 0-based CRS in the caller's order, columns ascending.
int brick_colour_order(int n, const int *row, const int *col, const int *dia, int edge, int cap,
                       std::vector<int> &perm, std::vector<int> &bstart,
                       std::vector<int> &bcolour);

void build_layout(int n, const int *row, const int *col, const int *dia, int spmv_rows,
                  int spmv_cap, int small_level, Layout &out, bool identity_s = false);

}  // namespace tfs
