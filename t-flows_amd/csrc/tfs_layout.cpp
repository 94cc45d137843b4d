// tfs_layout.cpp -- host-side construction of the S / L / U layouts.
#include "tfs_layout.hpp"

#include <algorithm>
#include <unordered_map>
#include <cstdlib>

namespace tfs {
namespace {

// The upper layout is the lower one backwards (see build_sweep).
constexpr bool kMirrorUpper = true;

// Breadth-first search from `start` over nodes whose mark differs from
// `stamp` (marking them); neighbours are taken in ascending degree, ties by
// number (Cuthill-McKee).  Appends the visit order to `out`, returns the
// number of BFS levels and, if asked, the nodes of the last level.
int cm_bfs(int start, int stamp, const std::vector<int> &row, const std::vector<int> &adj,
           const std::vector<int> &deg, std::vector<int> &mark, std::vector<int> &out,
           std::vector<int> *last_level) {
  size_t head = out.size(), level_begin = head, level_end = head + 1;
  out.push_back(start);
  mark[start] = stamp;
  int levels = 1;
  std::vector<int> nb;
  while (head < out.size()) {
    if (head == level_end) {
      level_begin = level_end;
      level_end = out.size();
      ++levels;
    }
    const int v = out[head++];
    nb.clear();
    for (int e = row[v]; e < row[v + 1]; ++e)
      if (mark[adj[e]] != stamp) {
        mark[adj[e]] = stamp;
        nb.push_back(adj[e]);
      }
    std::sort(nb.begin(), nb.end(), [&](int a, int b) {
      return deg[a] != deg[b] ? deg[a] < deg[b] : a < b;
    });
    out.insert(out.end(), nb.begin(), nb.end());
  }
  if (last_level) last_level->assign(out.begin() + (long)level_begin, out.end());
  return levels;
}

// Reverse Cuthill-McKee (Cuthill & McKee 1969; George 1971), starting each
// connected component from a pseudo-peripheral node (George & Liu 1979).
std::vector<int> rcm(int n, const int *row, const int *col) {
  // symmetric adjacency without the diagonal
  std::vector<int> cnt(n + 1, 0);
  for (int i = 0; i < n; ++i)
    for (int e = row[i]; e < row[i + 1]; ++e)
      if (col[e] != i) {
        cnt[i + 1]++;
        cnt[col[e] + 1]++;
      }
  for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  std::vector<int> adj(cnt[n]), fill(cnt.begin(), cnt.end() - 1);
  for (int i = 0; i < n; ++i)
    for (int e = row[i]; e < row[i + 1]; ++e)
      if (col[e] != i) {
        adj[fill[i]++] = col[e];
        adj[fill[col[e]]++] = i;
      }
  // drop duplicates (a structurally symmetric pattern lists each pair twice)
  std::vector<int> arow(n + 1, 0), a2;
  a2.reserve(adj.size() / 2 + 1);
  for (int i = 0; i < n; ++i) {
    std::sort(adj.begin() + cnt[i], adj.begin() + cnt[i + 1]);
    int prev = -1;
    for (int e = cnt[i]; e < cnt[i + 1]; ++e)
      if (adj[e] != prev) a2.push_back(prev = adj[e]);
    arow[i + 1] = (int)a2.size();
  }
  std::vector<int>().swap(adj);
  std::vector<int> deg(n);
  for (int i = 0; i < n; ++i) deg[i] = arow[i + 1] - arow[i];

  // mark: 0 = unvisited; 1 = placed in `order`; >= 2 = probe stamps
  std::vector<int> mark(n, 0), order, tmp, last;
  order.reserve(n);
  int stamp = 2;
  for (int s0 = 0; s0 < n; ++s0) {
    if (mark[s0] == 1) continue;
    int r = s0, ecc = -1;
    for (int it = 0; it < 4; ++it) {
      tmp.clear();
      // probe BFS inside the component (placed nodes carry mark 1 and are
      // in other components, so they are never reached)
      const int levels = cm_bfs(r, stamp++, arow, a2, deg, mark, tmp, &last);
      if (levels <= ecc) break;
      ecc = levels;
      int best = last[0];
      for (int v : last)
        if (deg[v] < deg[best] || (deg[v] == deg[best] && v < best)) best = v;
      r = best;
    }
    cm_bfs(r, 1, arow, a2, deg, mark, order, nullptr);
  }
  std::reverse(order.begin(), order.end());
  return order;  // order[k] = caller row placed at position k
}

// With `mirror` (the lower layout, for the upper triangle) the levels are
// the lower levels counted from the deepest one, ties in reverse S order: the
// upper order is then the lower order backwards.  That is a valid order for
// the backward sweep (a row's upper dependencies have larger caller numbers,
// hence deeper lower levels, hence earlier mirrored positions), has the same
// number of levels as its own level order, and lets the backward sweep read
// the forward sweep's output and re-arm it in contiguous runs.
void build_sweep(int n, const int *row, const int *col, const int *dia, bool lower,
                 const std::vector<int> &s_of_orig, int small_level, SweepLayout &S,
                 const SweepLayout *mirror = nullptr) {
  std::vector<int> lev(n, 0);
  int nlev = 0;
  if (mirror) {
    nlev = mirror->nlev;
    for (int l = 0; l < nlev; ++l)
      for (int p = mirror->lev_ptr[l]; p < mirror->lev_ptr[l + 1]; ++p)
        lev[mirror->orig[p]] = nlev - 1 - l;
  } else if (lower) {
    for (int i = 0; i < n; ++i) {
      int l = 0;
      for (int j = row[i]; j < dia[i]; ++j) l = std::max(l, lev[col[j]] + 1);
      lev[i] = l;
      nlev = std::max(nlev, l + 1);
    }
  } else {
    for (int i = n - 1; i >= 0; --i) {
      int l = 0;
      for (int j = dia[i] + 1; j < row[i + 1]; ++j) l = std::max(l, lev[col[j]] + 1);
      lev[i] = l;
      nlev = std::max(nlev, l + 1);
    }
  }
  S.nlev = nlev;
  S.lev_ptr.assign(nlev + 1, 0);
  for (int i = 0; i < n; ++i) S.lev_ptr[lev[i] + 1]++;
  for (int l = 0; l < nlev; ++l) S.lev_ptr[l + 1] += S.lev_ptr[l];
  // order: by level, ties in S order (walk caller rows in S order), or in
  // reverse S order for a mirrored layout
  std::vector<int> s_orig(n);
  for (int i = 0; i < n; ++i) s_orig[s_of_orig[i]] = i;
  std::vector<int> fill(S.lev_ptr.begin(), S.lev_ptr.end() - 1);
  S.orig.assign(n, 0);
  S.of_orig.assign(n, 0);
  for (int k = 0; k < n; ++k) {
    const int i = s_orig[mirror ? n - 1 - k : k];
    const int p = fill[lev[i]]++;
    S.orig[p] = i;
    S.of_orig[i] = p;
  }
  S.ptr.assign(n + 1, 0);
  for (int p = 0; p < n; ++p) {
    const int i = S.orig[p];
    S.ptr[p + 1] = S.ptr[p] + (lower ? dia[i] - row[i] : row[i + 1] - dia[i] - 1);
  }
  S.col.assign(S.ptr[n], 0);
  S.src.assign(S.ptr[n], 0);
  S.diag_src.assign(n, 0);
  for (int p = 0; p < n; ++p) {
    const int i = S.orig[p];
    const int j0 = lower ? row[i] : dia[i] + 1, j1 = lower ? dia[i] : row[i + 1];
    int e = S.ptr[p];
    for (int j = j0; j < j1; ++j, ++e) {
      S.col[e] = S.of_orig[col[j]];
      S.src[e] = j;
    }
    S.diag_src[p] = dia[i];
  }
  S.segs.clear();
  for (int l = 0; l < nlev; ++l) {
    const int rows = S.lev_ptr[l + 1] - S.lev_ptr[l];
    const bool small = rows <= small_level;
    if (small && !S.segs.empty() && S.segs.back().small) {
      S.segs.back().l1 = l + 1;
      S.segs.back().rows += rows;
    } else {
      S.segs.push_back({l, l + 1, rows, small});
    }
  }
}

}  // namespace

int colour_order(int n, const int *row, const int *col, std::vector<int> &perm,
                 std::vector<int> *colour_out, int block) {
  // symmetrised adjacency: the pattern may be unsymmetric
  std::vector<int> deg(n + 1, 0);
  for (int i = 0; i < n; ++i)
    for (int k = row[i]; k < row[i + 1]; ++k)
      if (col[k] != i) {
        deg[i + 1]++;
        deg[col[k] + 1]++;
      }
  for (int i = 0; i < n; ++i) deg[i + 1] += deg[i];
  std::vector<int> adj(deg[n]), fill(deg.begin(), deg.end() - 1);
  for (int i = 0; i < n; ++i)
    for (int k = row[i]; k < row[i + 1]; ++k)
      if (col[k] != i) {
        adj[fill[i]++] = col[k];
        adj[fill[col[k]]++] = i;
      }
  // blocks of `block` consecutive rows are coloured as units: a block takes
  // the smallest colour that no neighbouring block has
  block = std::max(1, block);
  const int nb = (n + block - 1) / block;
  std::vector<int> bcol(nb, -1), seen;  // seen[c] == B: colour c taken by a neighbour of B
  int ncol = 0;
  for (int B = 0; B < nb; ++B) {
    const int i0 = B * block, i1 = std::min(n, i0 + block);
    for (int i = i0; i < i1; ++i)
      for (int k = deg[i]; k < deg[i + 1]; ++k) {
        const int Bj = adj[k] / block;
        if (Bj != B && bcol[Bj] >= 0) seen[bcol[Bj]] = B;
      }
    int c = 0;
    while (c < ncol && seen[c] == B) ++c;
    if (c == ncol) {
      ++ncol;
      seen.push_back(-1);
    }
    bcol[B] = c;
  }
  std::vector<int> colour(n);
  for (int i = 0; i < n; ++i) colour[i] = bcol[i / block];
  std::vector<int> start(ncol + 1, 0);
  for (int i = 0; i < n; ++i) start[colour[i] + 1]++;
  for (int c = 0; c < ncol; ++c) start[c + 1] += start[c];
  perm.assign(n, 0);
  for (int i = 0; i < n; ++i) perm[start[colour[i]]++] = i;
  if (colour_out) *colour_out = std::move(colour);
  return ncol;
}

int block_colour_order(int n, const int *row, const int *col, int block, std::vector<int> &perm,
                       std::vector<int> *colour_out, std::vector<int> *block_out) {
  // symmetrised adjacency, each row's neighbours ascending and distinct
  std::vector<int> deg(n + 1, 0);
  for (int i = 0; i < n; ++i)
    for (int k = row[i]; k < row[i + 1]; ++k)
      if (col[k] != i) {
        deg[i + 1]++;
        deg[col[k] + 1]++;
      }
  for (int i = 0; i < n; ++i) deg[i + 1] += deg[i];
  std::vector<int> adj(deg[n]), fill(deg.begin(), deg.end() - 1);
  for (int i = 0; i < n; ++i)
    for (int k = row[i]; k < row[i + 1]; ++k)
      if (col[k] != i) {
        adj[fill[i]++] = col[k];
        adj[fill[col[k]]++] = i;
      }
  for (int i = 0; i < n; ++i) std::sort(adj.begin() + deg[i], adj.begin() + deg[i + 1]);
  block = std::max(1, block);
  // blocks grown breadth-first from the lowest unassigned vertex
  std::vector<int> blk(n, -1), order;  // order: vertices block by block, discovery order
  std::vector<int> bstart(1, 0);
  order.reserve(n);
  for (int s = 0; s < n; ++s) {
    if (blk[s] >= 0) continue;
    const int b = (int)bstart.size() - 1;
    size_t head = order.size();
    const size_t first = head;
    blk[s] = b;
    order.push_back(s);
    while (head < order.size() && (int)(order.size() - first) < block) {
      const int v = order[head++];
      for (int e = deg[v]; e < deg[v + 1] && (int)(order.size() - first) < block; ++e) {
        const int w = adj[e];
        if (blk[w] < 0) {
          blk[w] = b;
          order.push_back(w);
        }
      }
    }
    bstart.push_back((int)order.size());
  }
  const int nb = (int)bstart.size() - 1;
  // greedy colouring of the block graph, blocks in growth order
  std::vector<int> bcol(nb, -1), seen;
  int ncol = 0;
  for (int B = 0; B < nb; ++B) {
    for (int k = bstart[B]; k < bstart[B + 1]; ++k) {
      const int v = order[k];
      for (int e = deg[v]; e < deg[v + 1]; ++e) {
        const int Bj = blk[adj[e]];
        if (Bj != B && bcol[Bj] >= 0) seen[bcol[Bj]] = B;
      }
    }
    int c = 0;
    while (c < ncol && seen[c] == B) ++c;
    if (c == ncol) {
      ++ncol;
      seen.push_back(-1);
    }
    bcol[B] = c;
  }
  // colour, then block, then discovery order
  std::vector<int> cstart(ncol + 1, 0);
  for (int B = 0; B < nb; ++B) cstart[bcol[B] + 1] += bstart[B + 1] - bstart[B];
  for (int c = 0; c < ncol; ++c) cstart[c + 1] += cstart[c];
  perm.assign(n, 0);
  for (int B = 0; B < nb; ++B)
    for (int k = bstart[B]; k < bstart[B + 1]; ++k) perm[cstart[bcol[B]]++] = order[k];
  if (colour_out) {
    colour_out->assign(n, 0);
    for (int i = 0; i < n; ++i) (*colour_out)[i] = bcol[blk[i]];
  }
  if (block_out) *block_out = std::move(blk);
  return ncol;
}

int brick_colour_order(int n, const int *row, const int *col, const int *dia, int edge, int cap,
                       std::vector<int> &perm, std::vector<int> &bstart,
                       std::vector<int> &bcolour) {
  perm.clear();
  bstart.assign(1, 0);
  bcolour.clear();
  if (n <= 0) return 0;
  edge = std::max(1, edge);
  cap = std::max(1, cap);
  // classes of the dependencies j -> i by the binary order of i - j
  constexpr int kClasses = 3;
  std::vector<long long> hist(32, 0);
  auto order_of = [](unsigned d) { return 31 - __builtin_clz(d); };
  long long ntri = 0;
  for (int i = 0; i < n; ++i)
    for (int e = row[i]; e < dia[i]; ++e) {
      hist[order_of((unsigned)(i - col[e]))]++;
      ++ntri;
    }
  std::vector<int> cls(32, 0);
  {
    int c = 0;
    long long acc = 0;
    bool any = false;
    for (int o = 0; o < 32; ++o) {
      if (!hist[o]) continue;
      if (any && c < kClasses - 1 && acc * kClasses >= (long long)(c + 1) * ntri) ++c;
      cls[o] = c;
      acc += hist[o];
      any = true;
    }
  }
  // potentials, row by row in caller order (dependencies come first)
  std::vector<int> phi((size_t)n * kClasses, 0);
  for (int i = 0; i < n; ++i) {
    int *f = &phi[(size_t)i * kClasses];
    for (int e = row[i]; e < dia[i]; ++e) {
      const int j = col[e], c = cls[order_of((unsigned)(i - j))];
      const int *g = &phi[(size_t)j * kClasses];
      for (int k = 0; k < kClasses; ++k) f[k] = std::max(f[k], g[k] + (k == c ? 1 : 0));
    }
  }
  // groups: one id per distinct brick key, in order of their first row; a
  // group's rows in caller order, split into blocks of at most cap rows
  long long span = 1;
  for (size_t i = 0; i < phi.size(); ++i) span = std::max<long long>(span, phi[i] / edge + 1);
  std::unordered_map<long long, int> gid;
  gid.reserve((size_t)n / (size_t)std::max(1, edge * edge * edge) + 16);
  std::vector<int> grp(n);
  std::vector<int> gcount;
  for (int i = 0; i < n; ++i) {
    const int *f = &phi[(size_t)i * kClasses];
    const long long key = (f[0] / edge) + span * ((f[1] / edge) + span * (long long)(f[2] / edge));
    auto it = gid.find(key);
    if (it == gid.end()) {
      it = gid.emplace(key, (int)gcount.size()).first;
      gcount.push_back(0);
    }
    grp[i] = it->second;
    gcount[(size_t)it->second]++;
  }
  std::vector<int>().swap(phi);
  const int ng = (int)gcount.size();
  std::vector<int> gfirst((size_t)ng + 1, 0);  // blocks of each group
  for (int g = 0; g < ng; ++g) gfirst[(size_t)g + 1] = gfirst[(size_t)g] + (gcount[(size_t)g] + cap - 1) / cap;
  const int nb = gfirst[(size_t)ng];
  std::vector<int> blk(n), seen_in_group((size_t)ng, 0);
  for (int i = 0; i < n; ++i) {
    const int g = grp[i];
    blk[i] = gfirst[(size_t)g] + seen_in_group[(size_t)g]++ / cap;
  }
  // greedy colouring of the block graph (symmetrised pattern), blocks in id order
  std::vector<int> bstart_c((size_t)nb + 1, 0);
  for (int i = 0; i < n; ++i) bstart_c[(size_t)blk[i] + 1]++;
  for (int b = 0; b < nb; ++b) bstart_c[(size_t)b + 1] += bstart_c[(size_t)b];
  std::vector<int> rows_of((size_t)n), fill(bstart_c.begin(), bstart_c.end() - 1);
  for (int i = 0; i < n; ++i) rows_of[(size_t)fill[(size_t)blk[i]]++] = i;  // caller order within
  // upper neighbours too: a block must differ from every adjacent block
  std::vector<int> deg(n + 1, 0);
  for (int i = 0; i < n; ++i)
    for (int e = row[i]; e < row[i + 1]; ++e)
      if (col[e] != i) {
        deg[i + 1]++;
        deg[col[e] + 1]++;
      }
  for (int i = 0; i < n; ++i) deg[i + 1] += deg[i];
  std::vector<int> adj((size_t)deg[n]);
  {
    std::vector<int> f2(deg.begin(), deg.end() - 1);
    for (int i = 0; i < n; ++i)
      for (int e = row[i]; e < row[i + 1]; ++e)
        if (col[e] != i) {
          adj[(size_t)f2[i]++] = col[e];
          adj[(size_t)f2[col[e]]++] = i;
        }
  }
  std::vector<int> bcol((size_t)nb, -1), taken;
  int ncol = 0;
  for (int b = 0; b < nb; ++b) {
    for (int k = bstart_c[(size_t)b]; k < bstart_c[(size_t)b + 1]; ++k) {
      const int v = rows_of[(size_t)k];
      for (int e = deg[v]; e < deg[v + 1]; ++e) {
        const int bb = blk[(size_t)adj[(size_t)e]];
        if (bb != b && bcol[(size_t)bb] >= 0) taken[(size_t)bcol[(size_t)bb]] = b;
      }
    }
    int c = 0;
    while (c < ncol && taken[(size_t)c] == b) ++c;
    if (c == ncol) {
      ++ncol;
      taken.push_back(-1);
    }
    bcol[(size_t)b] = c;
  }
  // order: colour, block, caller order
  std::vector<int> cstart((size_t)ncol + 1, 0);
  for (int b = 0; b < nb; ++b) cstart[(size_t)bcol[(size_t)b] + 1]++;
  for (int c = 0; c < ncol; ++c) cstart[(size_t)c + 1] += cstart[(size_t)c];
  std::vector<int> border((size_t)nb);
  {
    std::vector<int> f3(cstart.begin(), cstart.end() - 1);
    for (int b = 0; b < nb; ++b) border[(size_t)f3[(size_t)bcol[(size_t)b]]++] = b;
  }
  perm.reserve(n);
  bstart.reserve((size_t)nb + 1);
  bcolour.reserve(nb);
  for (int k = 0; k < nb; ++k) {
    const int b = border[(size_t)k];
    for (int q = bstart_c[(size_t)b]; q < bstart_c[(size_t)b + 1]; ++q) perm.push_back(rows_of[(size_t)q]);
    bstart.push_back((int)perm.size());
    bcolour.push_back(bcol[(size_t)b]);
  }
  return ncol;
}

void build_layout(int n, const int *row, const int *col, const int *dia, int spmv_rows,
                  int spmv_cap, int small_level, Layout &out, bool identity_s) {
  out.n = n;
  out.nnz = n > 0 ? row[n] : 0;
  if (identity_s) {
    out.s_orig.resize(n);
    for (int i = 0; i < n; ++i) out.s_orig[i] = i;
  } else {
    out.s_orig = rcm(n, row, col);
  }
  out.s_of_orig.assign(n, 0);
  for (int k = 0; k < n; ++k) out.s_of_orig[out.s_orig[k]] = k;
  out.s_row.assign(n + 1, 0);
  for (int k = 0; k < n; ++k) {
    const int i = out.s_orig[k];
    out.s_row[k + 1] = out.s_row[k] + row[i + 1] - row[i];
  }
  out.s_col.assign(out.nnz, 0);
  out.s_src.assign(out.nnz, 0);
  out.s_diag_src.assign(n, 0);
  int bw = 0;
  for (int k = 0; k < n; ++k) {
    const int i = out.s_orig[k];
    int e = out.s_row[k];
    for (int j = row[i]; j < row[i + 1]; ++j, ++e) {
      out.s_col[e] = out.s_of_orig[col[j]];
      out.s_src[e] = j;
      bw = std::max(bw, std::abs(out.s_col[e] - k));
    }
    out.s_diag_src[k] = dia[i];
  }
  out.bandwidth = bw;
  // SpMV row blocks: <= spmv_rows rows and <= spmv_cap entries each
  out.s_rblk.assign(1, 0);
  for (int r = 0; r < n;) {
    int e = r + 1;
    while (e < n && e - r < spmv_rows && out.s_row[e + 1] - out.s_row[r] <= spmv_cap) ++e;
    out.s_rblk.push_back(e);
    r = e;
  }
  build_sweep(n, row, col, dia, true, out.s_of_orig, small_level, out.L);
  build_sweep(n, row, col, dia, false, out.s_of_orig, small_level, out.U,
              kMirrorUpper ? &out.L : nullptr);
  out.l_to_s.assign(n, 0);
  out.u_to_l.assign(n, 0);
  out.s_to_u.assign(n, 0);
  for (int p = 0; p < n; ++p) out.l_to_s[p] = out.s_of_orig[out.L.orig[p]];
  for (int q = 0; q < n; ++q) out.u_to_l[q] = out.L.of_orig[out.U.orig[q]];
  for (int k = 0; k < n; ++k) out.s_to_u[k] = out.U.of_orig[out.s_orig[k]];
}

}  // namespace tfs
