// tfs_kernels.hpp -- the solver's device kernels (gfx950, wave64).
//
// Value-carrying kernels are templates on the storage type T (double for the
// drop-in, float for the inner solves of iterative refinement); every dot
// product and norm accumulates in double whatever T is.
//
//   k_spmv          CSR SpMV over S rows, one wavefront per row block
//                   (<= 64 rows, <= kSpmvCap entries): col/val streamed with
//                   consecutive lanes on consecutive entries, products staged
//                   in LDS, each lane then sums its row left to right -- the
//                   reference's order (Cg.f90:128-134), so FP64 results are
//                   bitwise the reference's.  No MFMA: SpMV is not a dense
//                   contraction.
//   k_sweep_syncfree  factor / forward / backward sweeps (Prec_Form.f90:32-40,
//                   Prec_Solve.f90:46-67) in one persistent launch each, rows
//                   waiting only on their own dependencies (value-as-flag).
//   k_level, k_levels_small  the same sweeps with one launch per level (used
//                   where a cooperative launch is unavailable).
//   vector kernels  single passes writing per-workgroup partial sums; k_finish
//                   reduces them in a fixed order (deterministic, fence-free)
//                   and advances the device-resident solve state.
#pragma once

#include <hip/hip_runtime.h>

#include "tfs_device.hpp"

namespace tfs {

constexpr int kSpmvWaves = 8;            // waves per SpMV workgroup (default shape)
constexpr int kSpmvThreads = kSpmvWaves * kWave;
constexpr int kSpmvRows = 64;            // rows per SpMV row block (one per lane)
constexpr int kSpmvCap = 512;            // entries per row block staged in LDS
constexpr int kRedThreads = 256;         // threads of the vector kernels
constexpr int kMaxGrid = 2048;           // bound of every partial-sum grid
constexpr int kSweepGrid = 2048;         // workgroup cap of one coloured-sweep launch
constexpr int kMaxPartials = 32768;      // partial-sum slots per matrix
constexpr int kLevelThreads = 256;       // threads per block, one level launch
constexpr int kSmallThreads = 1024;      // one workgroup walking small levels
constexpr int kSmallLevel = 2048;        // a level this small joins a walked run
constexpr int kChunk = 256;              // threads per persistent-sweep workgroup
constexpr int kPrefetch = 8;             // row entries held in registers
constexpr unsigned kSpinLimit = 1u << 20;
constexpr int kSweepGroups = 768;        // cap on a persistent sweep's workgroups
constexpr int kSweepGroupsWide = 1536;   // the wider grid the schedule tuner also times
constexpr int kInnerBlock = 1;           // rows per colour block of the FP32 inner order

// EPI_AUX0..5 store the plain sum into aux[k] (dot products), EPI_NAUX0..5
// its square root (norms).
enum Epi {
  EPI_BNRM = 0, EPI_INIRES, EPI_RHO, EPI_ALPHA, EPI_NORM,
  EPI_AUX0, EPI_AUX1, EPI_AUX2, EPI_AUX3, EPI_AUX4, EPI_AUX5,
  EPI_NAUX0, EPI_NAUX1, EPI_NAUX2, EPI_NAUX3, EPI_NAUX4, EPI_NAUX5
};

// SpMV modes (k_spmv, k_ell64): plain y = A x; residual y = y - A x (term by
// term, Residual.f90:22-27); dot: y = A x and partial sum of w.y
// (Cg.f90:148-153, Cgs.f90:157-162); cgs tail: q2 = A p1, xs += alpha p1,
// rs -= alpha q2 and partial ||rs||^2 (Cgs.f90:182-221, row-local once alpha
// is known).
enum SpmvMode { SPMV_PLAIN = 0, SPMV_RESID, SPMV_DOT, SPMV_CGS_TAIL };

// Reductions that run whatever the solve's state (the rest return at once
// once the solve is done).
__device__ __forceinline__ bool epi_always(int epi) { return epi == EPI_BNRM || epi >= EPI_AUX0; }

// ---------------------------------------------------------------------------
// Solve-state transitions (Cg.f90:48-58,75-90,108-122,148-153,167-181).
__device__ inline void epilogue(int epi, double tot, KState *st) {
  switch (epi) {
    case EPI_BNRM: {
      const double b = sqrt(tot);
      st->bnrm2 = b;
      if (b < st->tol) {
        st->done = 1; st->niter_out = 0; st->fin_res = b; st->err = b;
      }
      break;
    }
    case EPI_INIRES: {
      const double e = sqrt(tot);
      st->err = e;
      st->ini_res = e;
      if (e < st->tol) {
        st->done = 1; st->niter_out = 0; st->fin_res = e;
      }
      break;
    }
    case EPI_RHO:
      st->rho = tot;
      if (st->iter + 1 > 1) st->beta = tot / st->rho_old;
      break;
    case EPI_ALPHA:
      st->pq = tot;
      st->alpha = st->rho / tot;
      break;
    case EPI_NORM: {
      const double e = sqrt(tot);
      const int it = st->iter + 1;
      st->iter = it;
      st->err = e;
      if (e < st->tol) {
        st->done = 1; st->niter_out = it; st->fin_res = e; st->status = kStOk;
      } else {
        st->rho_old = st->rho;
        if (!isfinite(e)) {
          st->done = 1; st->niter_out = st->cap + 1; st->fin_res = e; st->status = kStBreakdown;
        } else if (it >= st->cap) {
          st->done = 1; st->niter_out = st->cap + 1; st->fin_res = e;
          st->status = kStNotConverged;
        }
      }
      break;
    }
    case EPI_AUX0: case EPI_AUX1: case EPI_AUX2: case EPI_AUX3: case EPI_AUX4: case EPI_AUX5:
      st->aux[epi - EPI_AUX0] = tot;
      break;
    case EPI_NAUX0: case EPI_NAUX1: case EPI_NAUX2: case EPI_NAUX3: case EPI_NAUX4:
    case EPI_NAUX5:
      st->aux[epi - EPI_NAUX0] = sqrt(tot);
      break;
    default:
      break;
  }
}

template <int NT>
__device__ __forceinline__ void store_partial(double v, double *partials) {
  __shared__ double lds[NT / kWave];
  const double s = block_sum<NT>(v, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Sums the partials of the previous kernel in index order and applies the
// state transition (one workgroup; visibility comes from the kernel boundary).
__global__ void __launch_bounds__(kRedThreads)
k_finish(int nparts, const double *__restrict__ partials, int epi, KState *st) {
  if (!epi_always(epi) && solve_done(st)) return;
  __shared__ double lds[kRedThreads / kWave];
  double v = 0.0;
  for (int k = threadIdx.x; k < nparts; k += kRedThreads) v += partials[k];
  v = block_sum<kRedThreads>(v, lds);
  if (threadIdx.x == 0) epilogue(epi, v, st);
}

#define TFS_GRID_STRIDE(i, n) \
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < (n); i += gridDim.x * kRedThreads)

// ---------------------------------------------------------------------------
// Vector kernels (single pass each).

template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_sumsq(int n, const T *__restrict__ v, int epi, const KState *st, double *partials) {
  if (!epi_always(epi) && solve_done(st)) return;
  double s = 0.0;
  TFS_GRID_STRIDE(i, n) {
    const double a = v[i];
    s += a * a;
  }
  store_partial<kRedThreads>(s, partials);
}

// Plain dot product with per-workgroup partials (outer loop of the
// refinement; the host reads the result).
template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_dot(int n, const T *__restrict__ a, const T *__restrict__ b, double *partials) {
  double s = 0.0;
  TFS_GRID_STRIDE(i, n) s += (double)a[i] * (double)b[i];
  store_partial<kRedThreads>(s, partials);
}

// p = z + beta p
__global__ void __launch_bounds__(kRedThreads)
k_xpby(int n, const double *__restrict__ z, double *__restrict__ p, double beta) {
  TFS_GRID_STRIDE(i, n) p[i] = z[i] + beta * p[i];
}

// x += alpha p; r -= alpha q; partial ||r||^2
__global__ void __launch_bounds__(kRedThreads)
k_update_xr(int n, double *__restrict__ x, double *__restrict__ r, const double *__restrict__ p,
            const double *__restrict__ q, double alpha, double *partials) {
  double s = 0.0;
  TFS_GRID_STRIDE(i, n) {
    x[i] = x[i] + alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    s += ri * ri;
  }
  store_partial<kRedThreads>(s, partials);
}

// rho = (r, z) where z comes from: 0 -> z is r itself; 1 -> z = r / d
// (Prec_Solve.f90:35-38), written; 2 -> z = zU[s_to_u[i]] (the backward sweep
// result, gathered into S order), written; 3 -> z given.
template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_rho(int n, const T *__restrict__ r, T *__restrict__ z, const T *__restrict__ d,
      const T *__restrict__ zU, const int *__restrict__ s_to_u, int mode, const KState *st,
      double *partials) {
  if (solve_done(st)) return;
  double s = 0.0;
  TFS_GRID_STRIDE(i, n) {
    const T ri = r[i];
    T zi;
    if (mode == 0) {
      zi = ri;
    } else if (mode == 1) {
      zi = ri / d[i];
      z[i] = zi;
    } else if (mode == 2) {
      zi = zU[s_to_u[i]];
      z[i] = zi;
    } else if (mode == 4) {  // the U order is the S order backwards
      zi = zU[n - 1 - i];
      z[i] = zi;
    } else if (mode == 5) {  // the same, z not stored (k_pupdate_rev reads zU backwards)
      zi = zU[n - 1 - i];
    } else {
      zi = z[i];
    }
    s += (double)ri * (double)zi;
  }
  store_partial<kRedThreads>(s, partials);
}

// p = z (first iteration) or p = z + beta p  (Cg.f90:114-123)
template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_pupdate(int n, const T *__restrict__ z, T *__restrict__ p, const KState *st) {
  if (solve_done(st)) return;
  const bool first = st->iter == 0;
  const double beta = st->beta;
  TFS_GRID_STRIDE(i, n) p[i] = first ? z[i] : (T)((double)z[i] + beta * (double)p[i]);
}

// The same with z read from the backward sweep's output zU, whose order is
// the S order backwards: z[i] = zU[n-1-i].
template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_pupdate_rev(int n, const T *__restrict__ zU, T *__restrict__ p, const KState *st) {
  if (solve_done(st)) return;
  const bool first = st->iter == 0;
  const double beta = st->beta;
  TFS_GRID_STRIDE(i, n) {
    const T zi = zU[n - 1 - i];
    p[i] = first ? zi : (T)((double)zi + beta * (double)p[i]);
  }
}

// x += alpha p; r -= alpha q; partial ||r||^2  (Cg.f90:159-172)
template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_axpy_norm(int n, T *__restrict__ x, T *__restrict__ r, const T *__restrict__ p,
            const T *__restrict__ q, const KState *st, double *partials) {
  if (solve_done(st)) return;
  const T alpha = (T)st->alpha;
  double s = 0.0;
  TFS_GRID_STRIDE(i, n) {
    x[i] = x[i] + alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    s += (double)ri * (double)ri;
  }
  store_partial<kRedThreads>(s, partials);
}

// CGS: u1 = r1 + beta q1; u2 = u1 + beta (q1 + beta u2)  (Cgs.f90:115-126)
__global__ void __launch_bounds__(kRedThreads)
k_cgs_u(int n, const double *__restrict__ r1, const double *__restrict__ q1,
        double *__restrict__ u1, double *__restrict__ u2, const KState *st) {
  if (solve_done(st)) return;
  const bool first = st->iter == 0;
  const double beta = st->beta;
  TFS_GRID_STRIDE(i, n) {
    if (first) {
      u1[i] = r1[i];
      u2[i] = r1[i];
    } else {
      const double a = r1[i] + beta * q1[i];
      u1[i] = a;
      u2[i] = a + beta * (q1[i] + beta * u2[i]);
    }
  }
}

// CGS: q1 = u1 - alpha v2; w = u1 + q1  (Cgs.f90:167-176)
__global__ void __launch_bounds__(kRedThreads)
k_cgs_q(int n, const double *__restrict__ u1, const double *__restrict__ v2,
        double *__restrict__ q1, double *__restrict__ w, const KState *st) {
  if (solve_done(st)) return;
  const double alpha = st->alpha;
  TFS_GRID_STRIDE(i, n) {
    const double q = u1[i] - alpha * v2[i];
    q1[i] = q;
    w[i] = u1[i] + q;
  }
}

// out[i] = (TO) in[idx[i]] (idx may be null); skipped once the solve is done.
template <typename TO, typename TI>
__global__ void __launch_bounds__(kRedThreads)
k_gather(int n, const TI *__restrict__ in, const int *__restrict__ idx, TO *__restrict__ out,
         const KState *st) {
  if (st && solve_done(st)) return;
  TFS_GRID_STRIDE(i, n) out[i] = (TO)in[idx ? idx[i] : i];
}

// The values in every layout in one pass, one thread per S row: the caller
// row's entries are read once (contiguous, caller entry order) and written to
// the S copy, the two diagonal copies, and the strictly-lower / strictly-upper
// runs of the row in the L / U layouts (which keep the caller's entry order:
// the lower run is the entries before the diagonal, the upper run those
// after it).
template <typename T>
__device__ __forceinline__ void values_put(int j, T v, int e0, int dj, int l0, int u0, int k,
                                           int p, T *s_val, T *s_diag, T *l_val, T *u_val,
                                           T *l_diag) {
  s_val[e0 + j] = v;
  if (j < dj) {
    l_val[l0 + j] = v;
  } else if (j == dj) {
    s_diag[k] = v;
    l_diag[p] = v;
  } else {
    u_val[u0 + (j - dj - 1)] = v;
  }
}

template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_values(int n, const double *__restrict__ val, const int *__restrict__ s_row,
         const int *__restrict__ s_src, const int *__restrict__ s_diag_src,
         const int *__restrict__ s_lpos, const int *__restrict__ s_to_u,
         const int *__restrict__ l_ptr, const int *__restrict__ u_ptr, T *__restrict__ s_val,
         T *__restrict__ s_diag, T *__restrict__ l_val, T *__restrict__ u_val,
         T *__restrict__ l_diag) {
  TFS_GRID_STRIDE(k, n) {
    const int e0 = s_row[k], len = s_row[k + 1] - e0;
    const int c0 = s_src[e0];           // the row's first entry in val[]
    const int dj = s_diag_src[k] - c0;  // the diagonal's place in the row
    const int p = s_lpos[k], q = s_to_u[k];
    const int l0 = l_ptr[p], u0 = u_ptr[q];
    // the first kPrefetch entries: all loads issued before any store
    T v[kPrefetch];
#pragma unroll
    for (int j = 0; j < kPrefetch; ++j)
      if (j < len) v[j] = (T)val[c0 + j];
#pragma unroll
    for (int j = 0; j < kPrefetch; ++j)
      if (j < len) values_put(j, v[j], e0, dj, l0, u0, k, p, s_val, s_diag, l_val, u_val, l_diag);
    for (int j = kPrefetch; j < len; ++j)
      values_put(j, (T)val[c0 + j], e0, dj, l0, u0, k, p, s_val, s_diag, l_val, u_val, l_diag);
  }
}

// out[idx[i]] = in[i]
template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_scatter(int n, const T *__restrict__ in, const int *__restrict__ idx, T *__restrict__ out) {
  TFS_GRID_STRIDE(i, n) out[idx[i]] = in[i];
}

// x = b / d (prec 1) or x = b (prec 0)
template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_diag_apply(int n, const T *__restrict__ b, const T *__restrict__ d, T *__restrict__ x, int prec,
             const KState *st) {
  if (st && solve_done(st)) return;
  TFS_GRID_STRIDE(i, n) x[i] = prec == 1 ? b[i] / d[i] : b[i];
}

template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_fill(int n, T *__restrict__ x, T v) {
  TFS_GRID_STRIDE(i, n) x[i] = v;
}

// Iterative refinement: x += e (FP64 update with the FP32 correction).
__global__ void __launch_bounds__(kRedThreads)
k_add_correction(int n, double *__restrict__ x, const float *__restrict__ e,
                 const int *__restrict__ idx) {
  TFS_GRID_STRIDE(i, n) {
    const int j = idx ? idx[i] : i;  // idx: a permutation, so no two lanes share a j
    x[j] = x[j] + (double)e[i];
  }
}

// ---------------------------------------------------------------------------
// FP32 inner operator of the reliable-update refinement, in difference form:
//   y_i = s_i x_i + sum_{j != i} a_ij (x_j - x_i),   s_i = sum_j a_ij,
// with a_ij rounded to FP32 and s_i summed in FP64 and then rounded.  The
// rounding of each product then scales with the local difference x_j - x_i
// rather than with x_j, so on smooth vectors -- the slow modes of a Krylov
// solve of a diffusion operator -- the FP32 operator's error stays at about
// eps32 relative to A x instead of eps32 * cond(A).  That is what lets an FP32
// Krylov process with FP64 residual replacement keep the FP64 iteration count
// on ill-conditioned Poisson systems (in the usual sum-of-products form the
// true residual stalls once eps32 * cond(A) ~ 1).
// Storage: sliced ELL, 64 rows per slice (one row per lane), the slice's width
// its longest off-diagonal run, entries column-major inside the slice (so
// every col/val load of a wave is one contiguous 256-byte run); padding
// entries carry the row's own column (or 0 past the last row) and value 0.
// Arithmetic is FP64 (free in a bandwidth-bound kernel); only storage is FP32.
constexpr int kEllWaves = 4;  // slices per workgroup (one wavefront each)

// Slice walk of the ELL kernels.  With `xcd` set, the 8 XCDs (workgroups are
// dispatched to them round-robin: XCD = blockIdx % 8) each walk one
// contiguous eighth of the slices, so the x gathers of one XCD's L2 fall in
// one window of the vector; otherwise workgroups take slices interleaved.
struct SliceWalk {
  int first, step, end;
  __device__ SliceWalk(int nslice, bool xcd) {
    const int wv = (int)threadIdx.x / kWave;
    if (xcd && gridDim.x >= 8) {
      const int x = (int)blockIdx.x % 8, k = (int)blockIdx.x / 8, g8 = (int)gridDim.x / 8;
      const int per = (nslice + 7) / 8;
      const int lo = x * per, hi = min(nslice, lo + per);
      // the last gridDim % 8 workgroups join XCDs 0.. (same interleave)
      first = lo + k * kEllWaves + wv;
      step = g8 * kEllWaves;
      end = k < g8 ? hi : lo;  // an extra workgroup has nothing of its own
    } else {
      first = (int)blockIdx.x * kEllWaves + wv;
      step = (int)gridDim.x * kEllWaves;
      end = nslice;
    }
  }
};
constexpr int kEllBatch = 8;  // entries per lane loaded before their gathers

template <bool DOT>
__global__ void __launch_bounds__(kEllWaves * kWave)
k_ell_diff(int n, int nslice, const long long *__restrict__ off, const int *__restrict__ col,
           const float *__restrict__ val, const float *__restrict__ rsum,
           const float *__restrict__ x, float *__restrict__ y, const KState *st,
           double *partials, int xcd) {
  if (st && solve_done(st)) return;
  const int lane = threadIdx.x % kWave;
  double red = 0.0;
  const SliceWalk sw(nslice, xcd != 0);
  for (int s = sw.first; s < sw.end; s += sw.step) {
    const int su = __builtin_amdgcn_readfirstlane(s);
    const long long e0 = off[su];
    const int w = (int)((off[su + 1] - e0) / kWave);
    const int i = su * kWave + lane;
    const bool live = i < n;
    const double xi = live ? (double)x[i] : 0.0;
    double acc = live ? (double)rsum[i] * xi : 0.0;
    const long long base = e0 + lane;
    for (int j0 = 0; j0 < w; j0 += kEllBatch) {
      int c[kEllBatch];
      float a[kEllBatch], xv[kEllBatch];
#pragma unroll
      for (int k = 0; k < kEllBatch; ++k)
        if (j0 + k < w) {
          c[k] = col[base + (long long)(j0 + k) * kWave];
          a[k] = val[base + (long long)(j0 + k) * kWave];
        }
#pragma unroll
      for (int k = 0; k < kEllBatch; ++k)
        if (j0 + k < w) xv[k] = x[c[k]];
#pragma unroll
      for (int k = 0; k < kEllBatch; ++k)
        if (j0 + k < w) acc = acc + (double)a[k] * ((double)xv[k] - xi);
    }
    if (live) {
      const float yi = (float)acc;
      y[i] = yi;
      if (DOT) red += xi * (double)yi;
    }
  }
  if (DOT) store_partial<kEllWaves * kWave>(red, partials);
}

// Incomplete-Cholesky preconditioner z = M^-1 r of a point-coloured inner
// order (rows of one colour are independent: every triangle entry of a row
// lies in another colour), over per-colour ELL slices (64 rows, entries
// column-major, as k_ell_diff), in FP64 arithmetic on FP32 storage:
//   forward   y_i = (r_i - sum_{j<i} a_ij y_j) d_i                (Prec_Solve.f90:46-53)
//   backward  z_i = (y_i / (d_i + TINY) - sum_{j>i} a_ij z_j) d_i   (:55-67, scale fused)
// The first colour has no lower entries (its forward pass is y = r d, a
// plain stream) and the last colour no upper ones: its z comes out of its
// forward launch (LAST), so red-black takes three launches.  (Forming the
// first colour's y where it is gathered instead of storing it doubles the
// gathers of the next colour: slower at 256^3.)  Every launch that writes z
// adds its rows' part of rho = (r, z) (Cg.f90:108-112) to its own partial
// sums.
template <bool BWD, bool LAST>
__global__ void __launch_bounds__(kEllWaves * kWave)
k_colour_sweep(int s0, int s1, const int *__restrict__ row0, const long long *__restrict__ off,
               const int *__restrict__ col, const float *__restrict__ val,
               const float *__restrict__ d, const float *__restrict__ r, float *__restrict__ y,
               float *__restrict__ z, int row_end, const KState *st, double *partials) {
  if (st && solve_done(st)) return;
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  const float *__restrict__ xin = BWD ? z : y;  // the dependencies
  double rho = 0.0;
  for (int s = s0 + blockIdx.x * kEllWaves + wv; s < s1; s += gridDim.x * kEllWaves) {
    const int su = __builtin_amdgcn_readfirstlane(s);
    const long long e0 = off[su];
    const int w = (int)((off[su + 1] - e0) / kWave);
    const int i = row0[su] + lane;
    const bool live = i < row_end;
    const double di = live ? (double)d[i] : 0.0;
    const double ri = live ? (double)r[i] : 0.0;
    double acc;
    if (BWD) {
      const double yi = live ? (double)y[i] : 0.0;
      acc = yi / (di + kTiny);
    } else {
      acc = ri;
    }
    const long long base = e0 + lane;
    for (int j0 = 0; j0 < w; j0 += kEllBatch) {
      int c[kEllBatch];
      float a[kEllBatch];
      double xv[kEllBatch];
#pragma unroll
      for (int k = 0; k < kEllBatch; ++k)
        if (j0 + k < w) {
          c[k] = col[base + (long long)(j0 + k) * kWave];
          a[k] = val[base + (long long)(j0 + k) * kWave];
        }
#pragma unroll
      for (int k = 0; k < kEllBatch; ++k)
        if (j0 + k < w) {
          const int cj = c[k];  // -1: padding (the slot may hold anything)
          xv[k] = cj < 0 ? 0.0 : (double)xin[cj];
        }
#pragma unroll
      for (int k = 0; k < kEllBatch; ++k)
        if (j0 + k < w) acc = acc - (double)a[k] * xv[k];
    }
    if (live) {
      const double out = acc * di;
      if (BWD || LAST) {
        const float zi = (float)(LAST && !BWD ? (out / (di + kTiny)) * di : out);
        z[i] = zi;
        rho += ri * (double)zi;
      } else {
        y[i] = (float)out;
      }
    }
  }
  if (BWD || LAST) store_partial<kEllWaves * kWave>(rho, partials);
}

// Sliced-ELL values (FP32 or FP64) from the FP64 caller values: src[e] is
// the caller entry of slot e, or -1 for padding.
template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_ell_values(long long len, const int *__restrict__ src, const double *__restrict__ val,
             T *__restrict__ eval) {
  for (long long e = (long long)blockIdx.x * kRedThreads + threadIdx.x; e < len;
       e += (long long)gridDim.x * kRedThreads) {
    const int k = src[e];
    eval[e] = k >= 0 ? (T)val[k] : T(0);
  }
}

// FP64 SpMV over a sliced ELL of all the entries (diagonal included), each
// row's entries in the caller's order (padding: column -1, skipped, so no
// term is added that the row does not have): one lane per row sums its
// products left to right, as k_spmv and the reference do (Cg.f90:128-134,
// Residual.f90:22-27) -- bitwise the same results -- but with every col/val
// load of a wave one contiguous 256-byte run and no row pointers or LDS
// staging.  Modes as k_spmv; the dot products are summed per workgroup.
template <int MODE>
__global__ void __launch_bounds__(kEllWaves * kWave)
k_ell64(int n, int nslice, const long long *__restrict__ off, const int *__restrict__ col,
        const double *__restrict__ val, const double *__restrict__ x, double *__restrict__ y,
        const double *__restrict__ w, double *__restrict__ xs, double *__restrict__ rs,
        const KState *st, double *partials, int xcd) {
  if (st && solve_done(st)) return;
  const int lane = threadIdx.x % kWave;
  const double alpha = MODE == SPMV_CGS_TAIL ? st->alpha : 0.0;
  double red = 0.0;
  const SliceWalk sw(nslice, xcd != 0);
  for (int s = sw.first; s < sw.end; s += sw.step) {
    const int su = __builtin_amdgcn_readfirstlane(s);
    const long long e0 = off[su];
    const int wd = (int)((off[su + 1] - e0) / kWave);
    const int i = su * kWave + lane;
    const bool live = i < n;
    double acc = MODE == SPMV_RESID && live ? y[i] : 0.0;
    const long long base = e0 + lane;
    for (int j0 = 0; j0 < wd; j0 += kEllBatch) {
      int c[kEllBatch];
      double a[kEllBatch], xv[kEllBatch];
#pragma unroll
      for (int k = 0; k < kEllBatch; ++k)
        if (j0 + k < wd) {
          c[k] = col[base + (long long)(j0 + k) * kWave];
          a[k] = val[base + (long long)(j0 + k) * kWave];
        }
#pragma unroll
      for (int k = 0; k < kEllBatch; ++k)
        if (j0 + k < wd) xv[k] = c[k] >= 0 ? x[c[k]] : 0.0;
#pragma unroll
      for (int k = 0; k < kEllBatch; ++k)
        if (j0 + k < wd && c[k] >= 0) {
          const double pr = a[k] * xv[k];
          acc = MODE == SPMV_RESID ? acc - pr : acc + pr;
        }
    }
    if (live) {
      y[i] = acc;
      if (MODE == SPMV_DOT) red += w[i] * acc;
      if (MODE == SPMV_CGS_TAIL) {
        xs[i] = xs[i] + alpha * x[i];
        const double ri = rs[i] - alpha * acc;
        rs[i] = ri;
        red += ri * ri;
      }
    }
  }
  if (MODE == SPMV_DOT || MODE == SPMV_CGS_TAIL)
    store_partial<kEllWaves * kWave>(red, partials);
}

// Row sums of the operator, summed in FP64 over the caller's entries of each
// S row, rounded once.
__global__ void __launch_bounds__(kRedThreads)
k_row_sums(int n, const int *__restrict__ s_row, const int *__restrict__ s_src,
           const double *__restrict__ val, float *__restrict__ rsum) {
  TFS_GRID_STRIDE(k, n) {
    double s = 0.0;
    for (int e = s_row[k]; e < s_row[k + 1]; ++e) s += val[s_src[e]];
    rsum[k] = (float)s;
  }
}

// Reliable-update iteration tail: c += alpha p (the FP32 direction folded into
// an FP64 correction every iteration -- an FP32 running correction would
// carry a rounding error that A amplifies by its largest eigenvalue),
// r -= alpha q (FP32 recurrence residual), partial ||r||^2.  With n0 > 0 it
// also does the next sweep's first step for the rows of colour 0 (the first
// n0 rows of a point-coloured order, which have no lower entries):
// y = r d, the same operations as that sweep launch, which is then skipped.
__global__ void __launch_bounds__(kRedThreads)
k_axpy_mixed(int n, double *__restrict__ c, float *__restrict__ r, const float *__restrict__ p,
             const float *__restrict__ q, const KState *st, double *partials, int n0,
             const float *__restrict__ d, float *__restrict__ y) {
  if (solve_done(st)) return;
  const double alpha = st->alpha;
  double s = 0.0;
  TFS_GRID_STRIDE(i, n) {
    c[i] = c[i] + alpha * (double)p[i];
    const float ri = (float)((double)r[i] - alpha * (double)q[i]);
    r[i] = ri;
    if (i < n0) y[i] = (float)((double)ri * (double)d[i]);
    s += (double)ri * (double)ri;
  }
  store_partial<kRedThreads>(s, partials);
}

// x[idx[i]] += c[i]; c[i] = 0 (the correction of one reliable-update segment
// moved into the solution; idx is a permutation or null).
__global__ void __launch_bounds__(kRedThreads)
k_fold_correction(int n, double *__restrict__ x, double *__restrict__ c,
                  const int *__restrict__ idx) {
  TFS_GRID_STRIDE(i, n) {
    const int j = idx ? idx[i] : i;
    x[j] = x[j] + c[i];
    c[i] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// Batched (block-diagonal) solves: many independent systems concatenated in
// one matrix.  Vector work is split into tiles -- contiguous S ranges inside
// one system -- one workgroup per tile; every system has its own KState, and
// a tile of a converged system does nothing.  Per-tile partial sums are
// reduced per system, in tile order, by k_finish_seg (one workgroup per
// system).  `all` is the batch-wide state: its `done` is set when every
// system is done, so whole launches return at once after that.
struct Tiles {
  const int *__restrict__ sys;
  const int *__restrict__ beg;
  const int *__restrict__ end;
};

#define TFS_TILE_LOOP(i, T_) \
  for (int i = (T_).beg[blockIdx.x] + threadIdx.x; i < (T_).end[blockIdx.x]; i += kRedThreads)

__device__ __forceinline__ bool tile_skip(const Tiles &tl, const KState *st, const KState *all) {
  if (solve_done(all)) return true;
  return __builtin_amdgcn_readfirstlane(st[tl.sys[blockIdx.x]].done) != 0;
}

template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_sumsq_seg(Tiles tl, const T *__restrict__ v, int epi, const KState *st, const KState *all,
            double *partials) {
  if (epi != EPI_BNRM && tile_skip(tl, st, all)) return;
  double s = 0.0;
  TFS_TILE_LOOP(i, tl) {
    const double a = v[i];
    s += a * a;
  }
  store_partial<kRedThreads>(s, partials);
}

template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_rho_seg(Tiles tl, const T *__restrict__ r, T *__restrict__ z, const T *__restrict__ d,
          const T *__restrict__ zU, const int *__restrict__ s_to_u, int mode, const KState *st,
          const KState *all, double *partials) {
  if (tile_skip(tl, st, all)) return;
  double s = 0.0;
  TFS_TILE_LOOP(i, tl) {
    const T ri = r[i];
    T zi;
    if (mode == 0) {
      zi = ri;
    } else if (mode == 1) {
      zi = ri / d[i];
      z[i] = zi;
    } else {
      zi = zU[s_to_u[i]];
      z[i] = zi;
    }
    s += (double)ri * (double)zi;
  }
  store_partial<kRedThreads>(s, partials);
}

template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_pupdate_seg(Tiles tl, const T *__restrict__ z, T *__restrict__ p, const KState *st,
              const KState *all) {
  if (tile_skip(tl, st, all)) return;
  const KState &S = st[tl.sys[blockIdx.x]];
  const bool first = S.iter == 0;
  const T beta = (T)S.beta;
  TFS_TILE_LOOP(i, tl) p[i] = first ? z[i] : z[i] + beta * p[i];
}

template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_dot_seg(Tiles tl, const T *__restrict__ a, const T *__restrict__ b, const KState *st,
          const KState *all, double *partials) {
  if (tile_skip(tl, st, all)) return;
  double s = 0.0;
  TFS_TILE_LOOP(i, tl) s += (double)a[i] * (double)b[i];
  store_partial<kRedThreads>(s, partials);
}

template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_axpy_norm_seg(Tiles tl, T *__restrict__ x, T *__restrict__ r, const T *__restrict__ p,
                const T *__restrict__ q, const KState *st, const KState *all, double *partials) {
  if (tile_skip(tl, st, all)) return;
  const T alpha = (T)st[tl.sys[blockIdx.x]].alpha;
  double s = 0.0;
  TFS_TILE_LOOP(i, tl) {
    x[i] = x[i] + alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    s += (double)ri * (double)ri;
  }
  store_partial<kRedThreads>(s, partials);
}

// The reliable-update iteration tail per system (k_axpy_mixed over tiles):
// c += alpha_s p (FP64 correction), r -= alpha_s q (FP32 recurrence residual),
// per-tile partial ||r||^2.
__global__ void __launch_bounds__(kRedThreads)
k_axpy_mixed_seg(Tiles tl, double *__restrict__ c, float *__restrict__ r,
                 const float *__restrict__ p, const float *__restrict__ q, const KState *st,
                 const KState *all, double *partials) {
  if (tile_skip(tl, st, all)) return;
  const double alpha = st[tl.sys[blockIdx.x]].alpha;
  double s = 0.0;
  TFS_TILE_LOOP(i, tl) {
    c[i] = c[i] + alpha * (double)p[i];
    const float ri = (float)((double)r[i] - alpha * (double)q[i]);
    r[i] = ri;
    s += (double)ri * (double)ri;
  }
  store_partial<kRedThreads>(s, partials);
}

// out = in on the tiles of systems still running (the batched residual is
// formed in a scratch vector and kept only where the solve continues).
template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_select_seg(Tiles tl, const T *__restrict__ in, T *__restrict__ out, const KState *st,
             const KState *all) {
  if (tile_skip(tl, st, all)) return;
  TFS_TILE_LOOP(i, tl) out[i] = in[i];
}

// One workgroup per system: fixed-order sum of the system's tile partials,
// the state transition, and the batch-wide done count.
__global__ void __launch_bounds__(kRedThreads)
k_finish_seg(const int *__restrict__ tile_first, const double *__restrict__ partials, int epi,
             KState *st, KState *all) {
  if (epi != EPI_BNRM && solve_done(all)) return;
  KState *S = st + blockIdx.x;
  if (epi != EPI_BNRM && __builtin_amdgcn_readfirstlane(S->done)) return;
  __shared__ double lds[kRedThreads / kWave];
  double v = 0.0;
  for (int k = tile_first[blockIdx.x] + threadIdx.x; k < tile_first[blockIdx.x + 1];
       k += kRedThreads)
    v += partials[k];
  v = block_sum<kRedThreads>(v, lds);
  if (threadIdx.x == 0) {
    const int was = S->done;
    epilogue(epi, v, S);
    if (!was && S->done) {
      const unsigned int before = atomicAdd(&all->ctr[0], 1u);
      if (before + 1 == (unsigned int)all->cap) all->done = 1;
    }
  }
}

// ---------------------------------------------------------------------------
// SpMV over S rows.  Modes: plain y = A x; residual y = y - A x (term by term,
// Residual.f90:22-27); dot: y = A x and partial sum of w.y (Cg.f90:148-153,
// Cgs.f90:157-162); cgs tail: q2 = A p1, xs += alpha p1, rs -= alpha q2 and
// partial ||rs||^2 (Cgs.f90:182-221, all row-local once alpha is known).

// WAVES: wavefronts per workgroup (2, 4 or 8; the row blocks, at most 64 rows
// each, are a runtime layout) -- the two shapes the SpMV tuner picks.
template <typename T, int MODE, int WAVES>
__global__ void __launch_bounds__(WAVES * kWave)
k_spmv(int nblk, const int *__restrict__ rblk, const int *__restrict__ row,
       const int *__restrict__ col, const T *__restrict__ val, const T *__restrict__ x,
       T *__restrict__ y, const T *__restrict__ w, T *__restrict__ xs, T *__restrict__ rs,
       const KState *st, double *partials) {
  if (st && solve_done(st)) return;
  __shared__ T prod[WAVES][kSpmvCap];
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  T *lds = prod[wv];
  const T alpha = MODE == SPMV_CGS_TAIL ? (T)st->alpha : T(0);
  double acc_red = 0.0;
  // Wavefront w of workgroup g takes row blocks g*WAVES + w, then every
  // gridDim*WAVES-th.  Measured against splits that give each XCD (each L2)
  // one contiguous range of blocks, this plain interleave moved more bytes
  // per second both from HBM (cache-flushed launches, 128^3 and 256^3) and
  // inside the solves.
  for (int b = blockIdx.x * WAVES + wv; b < nblk; b += gridDim.x * WAVES) {
    // b is the same in every lane: the block's bounds come through the
    // scalar unit
    const int bu = __builtin_amdgcn_readfirstlane(b);
    const int r0 = rblk[bu], r1 = rblk[bu + 1];
    const int e0 = row[r0], e1 = row[r1];
    T yr = T(0);
    int r = -1;
    if (e1 - e0 <= kSpmvCap) {
      // all of the block's col/val loads are issued before the first x
      // gather (kSpmvCap / kWave per lane), so each wave keeps up to 8
      // independent loads in flight instead of one dependent pair
      // (indices past the block are clamped to its last entry: every load is
      // unconditional, so none waits on another, and the clamped ones hit
      // the cache line already being read)
      constexpr int kPer = kSpmvCap / kWave;
      int ci[kPer];
      T vi[kPer];
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int e = min(e0 + lane + u * kWave, e1 - 1);
        ci[u] = col[e];
        vi[u] = val[e];
      }
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int e = e0 + lane + u * kWave;
        const T pr = vi[u] * x[ci[u]];
        if (e < e1) lds[e - e0] = pr;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (lane < r1 - r0) {
        r = r0 + lane;
        const int a = row[r] - e0, z = row[r + 1] - e0;
        T acc = MODE == SPMV_RESID ? y[r] : T(0);
        if (MODE == SPMV_RESID) {
          for (int k = a; k < z; ++k) acc = acc - lds[k];
        } else {
          for (int k = a; k < z; ++k) acc = acc + lds[k];
        }
        yr = acc;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
      // one row longer than the block capacity: the wave forms its products
      // kSpmvCap at a time, lane 0 adds them in the row's order (as short
      // rows are added, so the result is the same)
      T acc = MODE == SPMV_RESID ? y[r0] : T(0);
      for (int c0 = e0; c0 < e1; c0 += kSpmvCap) {
        const int c1 = min(e1, c0 + kSpmvCap);
        for (int e = c0 + lane; e < c1; e += kWave) lds[e - c0] = val[e] * x[col[e]];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane == 0) {
          if (MODE == SPMV_RESID) {
            for (int k = 0; k < c1 - c0; ++k) acc = acc - lds[k];
          } else {
            for (int k = 0; k < c1 - c0; ++k) acc = acc + lds[k];
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      if (lane == 0) {
        r = r0;
        yr = acc;
      }
    }
    if (r >= 0) {
      y[r] = yr;
      if (MODE == SPMV_DOT) acc_red += (double)w[r] * (double)yr;
      if (MODE == SPMV_CGS_TAIL) {
        xs[r] = xs[r] + alpha * x[r];
        const T ri = rs[r] - alpha * yr;
        rs[r] = ri;
        acc_red += (double)ri * (double)ri;
      }
    }
  }
  if (MODE == SPMV_DOT || MODE == SPMV_CGS_TAIL) store_partial<WAVES * kWave>(acc_red, partials);
}

// ---------------------------------------------------------------------------
// The diagonal incomplete Cholesky over a level-contiguous layout.
//   ptr/col/val : the triangle entries of each position (caller's order),
//                 col in this layout's positions;
//   factor : d[p] = 1 / (diag[p] - sum d[col] val val)        (Prec_Form.f90:32-40)
//   fwd    : x[p] = (b[in_idx[p]] - sum val x[col]) d[p]       (Prec_Solve.f90:46-53)
//   bwd    : x[p] = (b[in_idx[p]] / (d[p] + TINY) - sum val x[col]) d[p]
//            -- the scale pass (:55-57) fused into the backward sweep (:60-67).
template <typename T>
struct Tri {
  const int *__restrict__ ptr;
  const int *__restrict__ col;
  const T *__restrict__ val;
  const T *__restrict__ diag;     // factor only
  const int *__restrict__ in_idx; // sweeps: gather index of the right-hand side; null
  int in_rev;                     // for the identity (in_rev 0) or the reversal p -> n-1-p
};                                // (in_rev n), formed instead of loaded

template <typename T>
__device__ __forceinline__ int tri_src(const Tri<T> &Tr, int p) {
  return Tr.in_idx ? Tr.in_idx[p] : Tr.in_rev ? Tr.in_rev - 1 - p : p;
}

enum SweepOp { OP_FACTOR = 0, OP_FWD, OP_BWD };

// One term of a row's sum, in FP64 whatever the storage type: the factor
// subtracts d_j a_ij^2 (Prec_Form.f90:36), the sweeps a_ij x_j
// (Prec_Solve.f90:49,63).  For FP64 storage this is the reference's own
// arithmetic; FP32 storage (the refinement's inner solves) is widened here,
// which keeps the inner preconditioner a fixed linear operator to FP32
// storage rounding -- with FP32 arithmetic the sweeps' rounding made the FP32
// CG recurrence lose up to 40 % more iterations than FP64 on red-black order.
template <int OP, typename T>
__device__ __forceinline__ double sweep_term(double s, T a, T v) {
  const double av = (double)a, vv = (double)v;
  return OP == OP_FACTOR ? s - vv * av * av : s - av * vv;
}

// One row of a level (all its dependencies are in earlier levels).  Three
// memory rounds: the row pointers and what the position indexes; the first
// kPrefetch entries and the right-hand side; their dependencies, all at
// once.  The terms are then folded in entry order (entries beyond the
// prefetch one by one), as the reference sums them.
template <typename T, int OP>
__device__ __forceinline__ void sweep_row(const Tri<T> &Tr, T *d, const T *b, T *x, int p) {
  T *out = OP == OP_FACTOR ? d : x;
  const int e0 = Tr.ptr[p], e1 = Tr.ptr[p + 1];
  double s, dp = 0.0;  // FP64 arithmetic whatever the storage type
  int src = 0;
  if (OP == OP_FACTOR) {
    s = Tr.diag[p];
  } else {
    dp = d[p];
    src = tri_src(Tr, p);
  }
  int c[kPrefetch];
  T a[kPrefetch], v[kPrefetch];
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k) {
    if (e0 + k < e1) {
      c[k] = Tr.col[e0 + k];
      a[k] = Tr.val[e0 + k];
    }
  }
  if (OP != OP_FACTOR) s = b[src];
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k)
    if (e0 + k < e1) v[k] = out[c[k]];
  if (OP == OP_BWD) s = s / (dp + kTiny);
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k)
    if (e0 + k < e1) s = sweep_term<OP>(s, a[k], v[k]);
  for (int e = e0 + kPrefetch; e < e1; ++e) s = sweep_term<OP>(s, Tr.val[e], out[Tr.col[e]]);
  out[p] = (T)(OP == OP_FACTOR ? 1.0 / s : s * dp);
}

template <typename T, int OP>
__global__ void __launch_bounds__(kLevelThreads)
k_level(int p0, int p1, Tri<T> Tr, T *d, const T *__restrict__ b, T *x, const KState *st) {
  if (st && solve_done(st)) return;
  const int p = p0 + blockIdx.x * kLevelThreads + threadIdx.x;
  if (p < p1) sweep_row<T, OP>(Tr, d, b, x, p);
}

// A run of small levels [l0, l1) walked by one workgroup; writes of one level
// are visible to the whole workgroup after the barrier.
template <typename T, int OP>
__global__ void __launch_bounds__(kSmallThreads)
k_levels_small(int l0, int l1, const int *__restrict__ lev_ptr, Tri<T> Tr, T *d,
               const T *__restrict__ b, T *x, const KState *st) {
  if (st && solve_done(st)) return;
  for (int l = l0; l < l1; ++l) {
    const int p1 = lev_ptr[l + 1];
    for (int p = lev_ptr[l] + threadIdx.x; p < p1; p += kSmallThreads)
      sweep_row<T, OP>(Tr, d, b, x, p);
    __syncthreads();
  }
}

// Sentinels: signalling NaNs, which no arithmetic produces (results of
// arithmetic on NaNs are quiet NaNs).
template <typename T> struct Sentinel;
template <> struct Sentinel<double> {
  static constexpr unsigned long long bits = 0xFFF4DEADBEEFC0DEull;
  __device__ static double value() { return __longlong_as_double((long long)bits); }
  __device__ static bool is(double v) { return (unsigned long long)__double_as_longlong(v) == bits; }
};
template <> struct Sentinel<float> {
  static constexpr unsigned int bits = 0xFFB1DEADu;
  __device__ static float value() { return __int_as_float((int)bits); }
  __device__ static bool is(float v) { return (unsigned int)__float_as_int(v) == bits; }
};

template <typename T>
__device__ __forceinline__ T coherent_load(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void coherent_store(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One row of a sweep, as a thread holds it.
template <typename T>
struct SweepRow {
  int p, e0, len, src, used;
  bool done;
  double s, dp;  // FP64 arithmetic whatever the storage type
  int pc[kPrefetch];  // the first kPrefetch entries: column positions ...
  T pv[kPrefetch];    // ... and values
};

// Memory round 1: what the position alone indexes.
template <typename T, int OP>
__device__ __forceinline__ void sweep_load1(SweepRow<T> &r, int p, int n, const Tri<T> &Tr,
                                            const T *d) {
  r.p = p;
  r.done = p >= n;
  r.e0 = r.len = r.used = r.src = 0;
  r.s = r.dp = 0.0;
  if (r.done) return;
  r.e0 = Tr.ptr[p];
  r.len = Tr.ptr[p + 1] - r.e0;
  if (OP == OP_FACTOR) {
    r.s = Tr.diag[p];
  } else {
    r.dp = d[p];
    r.src = tri_src(Tr, p);
  }
}

// Memory round 2: the row's entries and its right-hand side, with the
// sentinel upkeep (FWD arms the backward sweep's output at p; BWD re-arms the
// forward sweep's output at the entry it consumes -- this sweep is that
// entry's only reader).
template <typename T, int OP>
__device__ __forceinline__ void sweep_load2(SweepRow<T> &r, const Tri<T> &Tr, T *b, T *arm) {
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k) {
    if (k < r.len) {
      r.pc[k] = Tr.col[r.e0 + k];
      r.pv[k] = Tr.val[r.e0 + k];
    }
  }
  if (r.done || OP == OP_FACTOR) return;
  r.s = b[r.src];
  if (OP == OP_FWD) arm[r.p] = Sentinel<T>::value();
  if (OP == OP_BWD) b[r.src] = Sentinel<T>::value();
}

// One pass over the row's outstanding dependencies: every prefetched one is
// loaded at once (plain loads on the first pass, device-coherent polls after
// it), then they are folded in entry order up to the first sentinel (the
// values after it are loaded again in the next pass); entries beyond the
// prefetch go one by one.  Stores the row's value once the last dependency
// is in.  Returns true when the row is done.
template <typename T, int OP>
__device__ __forceinline__ bool sweep_pass(SweepRow<T> &r, const Tri<T> &Tr, T *out, bool first) {
  if (r.done) return true;
  T v[kPrefetch];
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k)
    if (k >= r.used && k < r.len) v[k] = first ? out[r.pc[k]] : coherent_load(&out[r.pc[k]]);
  bool blocked = false;
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k) {
    if (!blocked && k >= r.used && k < r.len) {
      if (Sentinel<T>::is(v[k])) {
        blocked = true;
      } else {
        r.s = sweep_term<OP>(r.s, r.pv[k], v[k]);
        r.used = k + 1;
      }
    }
  }
  for (int e = r.e0 + r.used; !blocked && e < r.e0 + r.len; ++e) {  // beyond the prefetch
    const T w = first ? out[Tr.col[e]] : coherent_load(&out[Tr.col[e]]);
    if (Sentinel<T>::is(w)) {
      blocked = true;
    } else {
      r.s = sweep_term<OP>(r.s, Tr.val[e], w);
      r.used = e - r.e0 + 1;
    }
  }
  if (blocked) return false;
  coherent_store(&out[r.p], (T)(OP == OP_FACTOR ? 1.0 / r.s : r.s * r.dp));
  r.done = true;
  return true;
}

template <typename T, int OP>
__device__ __forceinline__ void sweep_give_up(SweepRow<T> &r, T *out, int *err) {
  coherent_store(err, 1);
  if (!r.done) coherent_store(&out[r.p], (T)(OP == OP_FACTOR ? 1.0 / r.s : r.s * r.dp));
  r.done = true;
}

// The polling loop of one row.  The row's value is stored INSIDE the loop,
// in the pass in which its last dependency arrives: a chunk can hold two
// levels, so a row may wait on another lane of its own wavefront.  The loop
// therefore runs until every lane of the wave is done (a wave vote), with
// finished lanes idling through it: a loop that a lane left on its own would
// have its exit path -- and a store placed on that path by the compiler --
// run only once every lane had left, which deadlocks on such a wait.
// The first pass reads with plain (cached) loads: within a sweep a location
// only ever changes from the sentinel to its final value, and every sentinel
// was stored by an earlier launch, so a plain load returns either the final
// value or a (possibly stale) sentinel; only a sentinel sends the row to the
// device-coherent polls, which bypass the XCD's L2.
template <typename T, int OP>
__device__ __forceinline__ void sweep_row(SweepRow<T> &r, const Tri<T> &Tr, T *out, int *err) {
  if (OP == OP_BWD && !r.done) r.s = r.s / (r.dp + kTiny);
  unsigned int spins = 0;
  sweep_pass<T, OP>(r, Tr, out, true);
  while (__any(!r.done)) {
    if (!r.done &&
        (++spins > kSpinLimit || ((spins & 255u) == 0 && coherent_load(err) != 0))) {
      // out of patience, or another row already gave up: record the error
      // and store a value so that the dependants finish too
      sweep_give_up<T, OP>(r, out, err);
    }
    __builtin_amdgcn_s_sleep(2);
    sweep_pass<T, OP>(r, Tr, out, false);
  }
}

// One persistent launch per sweep, synchronisation-free in the manner of
// Liu et al., "A Synchronization-Free Algorithm for Parallel Sparse
// Triangular Solves" (Euro-Par 2016), with the computed value itself as the
// ready flag: the output array holds a sentinel until a row's result is
// stored.  Each row waits only on its own dependencies, polling them with
// device-coherent loads and folding them into its sum strictly in the
// caller's entry order -- so the result is bitwise the level-scheduled one.
// Flag and data share one location, so no fence or counter is needed.  One
// row per thread; a chunk is kChunk consecutive positions (level order).
// Every wait is bounded (kSpinLimit polls): on expiry the kernel records an
// error and stores a value, so it always terminates.
//
// DYN selects the schedule.  false -- static: workgroup w takes chunks w,
// w+G, ...; positions depend only on lower positions, and with all G
// workgroups resident (G is capped at the kernel's occupancy) the lowest
// unfinished row always belongs to a running workgroup whose dependencies
// are complete.  (Issuing the next chunk's loads while the current rows
// wait -- software pipelining -- measured no faster: 128^3 sweeps 75.7 us
// against 73.2, 256^3 equal; on the 62-level FP32 brick copy of a 256^3 box
// 389 us against 249, and the 766-level FP64 sweep 1452 against 1435 us.)
// true -- dynamic: chunks are taken in order from a counter (ctr[0]), so a
// chunk is only ever held by a running workgroup and the schedule is
// deadlock-free whatever else shares the GPU (used when several solves run
// concurrently).  The last workgroup to leave (ctr[1]) resets both counters
// for the next launch.
template <typename T, int OP, bool DYN>
__global__ void __launch_bounds__(kChunk)
k_sweep_syncfree(int n, Tri<T> Tr, T *d, T *b, T *x, T *arm, const KState *st, int *err,
                 unsigned int *ctr) {
  if (st && solve_done(st)) return;
  T *out = OP == OP_FACTOR ? d : x;
  const int nchunks = (n + kChunk - 1) / kChunk;
  const int tid = (int)threadIdx.x;
  if (!DYN) {
    const int G = (int)gridDim.x;
    int c = (int)blockIdx.x;
    for (; c < nchunks; c += G) {
      SweepRow<T> r;
      sweep_load1<T, OP>(r, c * kChunk + tid, n, Tr, d);
      sweep_load2<T, OP>(r, Tr, b, arm);
      sweep_row<T, OP>(r, Tr, out, err);
    }
    return;
  }
  __shared__ int s_chunk;
  for (;;) {
    __syncthreads();  // everyone has read the previous s_chunk
    if (tid == 0) s_chunk = (int)atomicAdd(&ctr[0], 1u);
    __syncthreads();
    const int c = s_chunk;
    if (c >= nchunks) break;
    SweepRow<T> r;
    sweep_load1<T, OP>(r, c * kChunk + tid, n, Tr, d);
    sweep_load2<T, OP>(r, Tr, b, arm);
    sweep_row<T, OP>(r, Tr, out, err);
  }
  if (tid == 0) {
    const unsigned int prev =
        __hip_atomic_fetch_add(&ctr[1], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (prev == gridDim.x - 1) {  // everyone else has left
      ctr[0] = 0u;
      ctr[1] = 0u;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(kRedThreads)
k_arm(int n, T *__restrict__ x) {
  TFS_GRID_STRIDE(i, n) x[i] = Sentinel<T>::value();
}

}  // namespace tfs
