// tfs_solver.hip -- MI355X (gfx950) drop-in for T-FlowS's Krylov solvers.
//
// Data layout (tfs_layout.hpp): the matrix lives in three row orders.  S
// (reverse Cuthill-McKee) carries the SpMV and every Krylov vector; L and U
// carry the lower / upper triangles level by level for the factor and the
// two sweeps.  Rows keep the caller's entry order, so every row sum is the
// reference's; the caller's x and b are permuted in and out once per solve.
//
// Kernels (all HBM-bound; none is a dense contraction, so no MFMA):
//   k_spmv          CSR SpMV over S rows, one wavefront per row block
//                   (<= 64 rows, <= kSpmvCap entries): col/val streamed with
//                   consecutive lanes on consecutive entries, products staged
//                   in LDS, each lane then sums its row left to right
//                   (Cg.f90:128-134) -- bitwise the reference's row sums.
//   k_level /       factor, forward and backward sweeps (Prec_Form.f90:32-40,
//   k_levels_small  Prec_Solve.f90:46-67) over the level-contiguous L / U
//                   layouts: one launch per big level, one workgroup walking
//                   each run of small levels.  The forward sweep gathers its
//                   right-hand side from the S vector, the backward sweep
//                   from the L result, so no separate permutation pass runs.
//   vector kernels  single-pass updates that write per-workgroup partial sums;
//                   a one-workgroup k_finish reduces them in a fixed order and
//                   advances the solve state, so reductions are deterministic
//                   and need no device-wide fences.  The solve state (KState)
//                   stays on the device; every kernel returns at once after
//                   convergence, so iterations are queued without host syncs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "tfs_device.hpp"
#include "tfs_layout.hpp"
#include "tfs_solver.h"

using tfs::KState;

namespace {

constexpr int kSpmvWaves = 4;            // waves per SpMV workgroup
constexpr int kSpmvThreads = kSpmvWaves * tfs::kWave;
constexpr int kSpmvRows = 64;            // rows per SpMV row block (one per lane)
constexpr int kSpmvCap = 512;            // entries per row block staged in LDS
constexpr int kRedThreads = 256;         // threads of the vector kernels
constexpr int kMaxGrid = 2048;           // bound of every partial-sum grid
constexpr int kLevelThreads = 256;       // threads per block, one level launch
constexpr int kSmallThreads = 1024;      // one workgroup walking small levels
constexpr int kSmallLevel = 2048;        // a level this small joins a walked run

enum Epi { EPI_BNRM = 0, EPI_INIRES, EPI_RHO, EPI_ALPHA, EPI_NORM };

// ---------------------------------------------------------------------------
// Solve-state transitions (Cg.f90:48-58,75-90,108-122,148-153,167-181).
__device__ void epilogue(int epi, double tot, KState *st) {
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
        st->done = 1; st->niter_out = it; st->fin_res = e; st->status = tfs::kStOk;
      } else {
        st->rho_old = st->rho;
        if (!isfinite(e)) {
          st->done = 1; st->niter_out = st->cap + 1; st->fin_res = e;
          st->status = tfs::kStBreakdown;
        } else if (it >= st->cap) {
          st->done = 1; st->niter_out = st->cap + 1; st->fin_res = e;
          st->status = tfs::kStNotConverged;
        }
      }
      break;
    }
    default:
      break;
  }
}

template <int NT>
__device__ __forceinline__ void store_partial(double v, double *partials) {
  __shared__ double lds[NT / tfs::kWave];
  const double s = tfs::block_sum<NT>(v, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Sums the partials of the previous kernel in index order and applies the
// state transition (one workgroup; visibility comes from the kernel boundary).
__global__ void __launch_bounds__(kRedThreads)
k_finish(int nparts, const double *__restrict__ partials, int epi, KState *st) {
  if (epi != EPI_BNRM && tfs::solve_done(st)) return;
  __shared__ double lds[kRedThreads / tfs::kWave];
  double v = 0.0;
  for (int k = threadIdx.x; k < nparts; k += kRedThreads) v += partials[k];
  v = tfs::block_sum<kRedThreads>(v, lds);
  if (threadIdx.x == 0) epilogue(epi, v, st);
}

#define GRID_STRIDE(i, n) \
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < (n); i += gridDim.x * kRedThreads)

// ---------------------------------------------------------------------------
// Vector kernels (single pass each).

__global__ void __launch_bounds__(kRedThreads)
k_sumsq(int n, const double *__restrict__ v, int epi, const KState *st, double *partials) {
  if (epi != EPI_BNRM && tfs::solve_done(st)) return;
  double s = 0.0;
  GRID_STRIDE(i, n) s += v[i] * v[i];
  store_partial<kRedThreads>(s, partials);
}

// rho = (r, z) where z comes from: 0 -> z is r itself; 1 -> z = r / d
// (Prec_Solve.f90:35-38), written; 2 -> z = zU[s_to_u[i]] (the backward sweep
// result, gathered into S order), written; 3 -> z given.
__global__ void __launch_bounds__(kRedThreads)
k_rho(int n, const double *__restrict__ r, double *__restrict__ z, const double *__restrict__ d,
      const double *__restrict__ zU, const int *__restrict__ s_to_u, int mode, const KState *st,
      double *partials) {
  if (tfs::solve_done(st)) return;
  double s = 0.0;
  GRID_STRIDE(i, n) {
    const double ri = r[i];
    double zi;
    if (mode == 0) {
      zi = ri;
    } else if (mode == 1) {
      zi = ri / d[i];
      z[i] = zi;
    } else if (mode == 2) {
      zi = zU[s_to_u[i]];
      z[i] = zi;
    } else {
      zi = z[i];
    }
    s += ri * zi;
  }
  store_partial<kRedThreads>(s, partials);
}

// p = z (first iteration) or p = z + beta p  (Cg.f90:114-123)
__global__ void __launch_bounds__(kRedThreads)
k_pupdate(int n, const double *__restrict__ z, double *__restrict__ p, const KState *st) {
  if (tfs::solve_done(st)) return;
  const bool first = st->iter == 0;
  const double beta = st->beta;
  GRID_STRIDE(i, n) p[i] = first ? z[i] : z[i] + beta * p[i];
}

// x += alpha p; r -= alpha q; partial ||r||^2  (Cg.f90:159-172)
__global__ void __launch_bounds__(kRedThreads)
k_axpy_norm(int n, double *__restrict__ x, double *__restrict__ r, const double *__restrict__ p,
            const double *__restrict__ q, const KState *st, double *partials) {
  if (tfs::solve_done(st)) return;
  const double alpha = st->alpha;
  double s = 0.0;
  GRID_STRIDE(i, n) {
    x[i] = x[i] + alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    s += ri * ri;
  }
  store_partial<kRedThreads>(s, partials);
}

// CGS: u1 = r1 + beta q1; u2 = u1 + beta (q1 + beta u2)  (Cgs.f90:115-126)
__global__ void __launch_bounds__(kRedThreads)
k_cgs_u(int n, const double *__restrict__ r1, const double *__restrict__ q1,
        double *__restrict__ u1, double *__restrict__ u2, const KState *st) {
  if (tfs::solve_done(st)) return;
  const bool first = st->iter == 0;
  const double beta = st->beta;
  GRID_STRIDE(i, n) {
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
  if (tfs::solve_done(st)) return;
  const double alpha = st->alpha;
  GRID_STRIDE(i, n) {
    const double q = u1[i] - alpha * v2[i];
    q1[i] = q;
    w[i] = u1[i] + q;
  }
}

// out[i] = in[idx[i]] (idx may be null: plain copy); skipped once done.
__global__ void __launch_bounds__(kRedThreads)
k_gather(int n, const double *__restrict__ in, const int *__restrict__ idx,
         double *__restrict__ out, const KState *st) {
  if (st && tfs::solve_done(st)) return;
  GRID_STRIDE(i, n) out[i] = in[idx ? idx[i] : i];
}

// out[idx[i]] = in[i]
__global__ void __launch_bounds__(kRedThreads)
k_scatter(int n, const double *__restrict__ in, const int *__restrict__ idx,
          double *__restrict__ out) {
  GRID_STRIDE(i, n) out[idx[i]] = in[i];
}

// x = b / d (prec 1) or x = b (prec 0), S order
__global__ void __launch_bounds__(kRedThreads)
k_diag_apply(int n, const double *__restrict__ b, const double *__restrict__ d,
             double *__restrict__ x, int prec, const KState *st) {
  if (st && tfs::solve_done(st)) return;
  GRID_STRIDE(i, n) x[i] = prec == 1 ? b[i] / d[i] : b[i];
}

__global__ void __launch_bounds__(kRedThreads)
k_fill(int n, double *__restrict__ x, double v) {
  GRID_STRIDE(i, n) x[i] = v;
}

// ---------------------------------------------------------------------------
// SpMV over S rows.  Modes: plain y = A x; residual y = y - A x (term by term,
// Residual.f90:22-27); dot: y = A x and partial sum of w.y (Cg.f90:148-153,
// Cgs.f90:157-162); cgs tail: q2 = A p1, xs += alpha p1, rs -= alpha q2 and
// partial ||rs||^2 (Cgs.f90:182-221, all row-local once alpha is known).
enum SpmvMode { SPMV_PLAIN = 0, SPMV_RESID, SPMV_DOT, SPMV_CGS_TAIL };

template <int MODE>
__global__ void __launch_bounds__(kSpmvThreads)
k_spmv(int nblk, const int *__restrict__ rblk, const int *__restrict__ row,
       const int *__restrict__ col, const double *__restrict__ val, const double *__restrict__ x,
       double *__restrict__ y, const double *__restrict__ w, double *__restrict__ xs,
       double *__restrict__ rs, const KState *st, double *partials) {
  if (st && tfs::solve_done(st)) return;
  __shared__ double prod[kSpmvWaves][kSpmvCap];
  const int lane = threadIdx.x % tfs::kWave, wv = threadIdx.x / tfs::kWave;
  double *lds = prod[wv];
  const double alpha = MODE == SPMV_CGS_TAIL ? st->alpha : 0.0;
  double acc_red = 0.0;
  for (int b = blockIdx.x * kSpmvWaves + wv; b < nblk; b += gridDim.x * kSpmvWaves) {
    const int r0 = rblk[b], r1 = rblk[b + 1];
    const int e0 = row[r0], e1 = row[r1];
    double yr = 0.0;
    int r = -1;
    if (e1 - e0 <= kSpmvCap) {
      for (int e = e0 + lane; e < e1; e += tfs::kWave) lds[e - e0] = val[e] * x[col[e]];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (lane < r1 - r0) {
        r = r0 + lane;
        const int a = row[r] - e0, z = row[r + 1] - e0;
        double acc = MODE == SPMV_RESID ? y[r] : 0.0;
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
      // one long row for the whole wave (tree order; rows this long do not
      // occur in finite-volume stencils)
      double acc = 0.0;
      for (int e = e0 + lane; e < e1; e += tfs::kWave) acc += val[e] * x[col[e]];
      acc = tfs::wave_sum(acc);
      if (lane == 0) {
        r = r0;
        yr = MODE == SPMV_RESID ? y[r] - acc : acc;
      }
    }
    if (r >= 0) {
      y[r] = yr;
      if (MODE == SPMV_DOT) acc_red += w[r] * yr;
      if (MODE == SPMV_CGS_TAIL) {
        xs[r] = xs[r] + alpha * x[r];
        const double ri = rs[r] - alpha * yr;
        rs[r] = ri;
        acc_red += ri * ri;
      }
    }
  }
  if (MODE == SPMV_DOT || MODE == SPMV_CGS_TAIL) store_partial<kSpmvThreads>(acc_red, partials);
}

// ---------------------------------------------------------------------------
// Level-scheduled factor and sweeps over a level-contiguous layout.
//   ptr/col/val : the triangle entries of each position (caller's order),
//                 col in this layout's positions;
//   factor : d[p] = 1 / (diag[p] - sum d[col] val val)        (Prec_Form.f90:32-40)
//   fwd    : x[p] = (b[in_idx[p]] - sum val x[col]) d[p]       (Prec_Solve.f90:46-53)
//   bwd    : x[p] = (b[in_idx[p]] / (d[p] + TINY) - sum val x[col]) d[p]
//            -- the scale pass (:55-57) fused into the backward sweep (:60-67).
struct Tri {
  const int *__restrict__ ptr;
  const int *__restrict__ col;
  const double *__restrict__ val;
  const double *__restrict__ diag;  // factor only
  const int *__restrict__ in_idx;   // sweeps: gather index of the right-hand side
};

enum SweepOp { OP_FACTOR = 0, OP_FWD, OP_BWD };

template <int OP>
__device__ __forceinline__ void sweep_row(const Tri &T, double *d, const double *b, double *x,
                                          int p) {
  const int e0 = T.ptr[p], e1 = T.ptr[p + 1];
  if (OP == OP_FACTOR) {
    double s = T.diag[p];
    for (int e = e0; e < e1; ++e) s = s - d[T.col[e]] * T.val[e] * T.val[e];
    d[p] = 1.0 / s;
  } else {
    const double dp = d[p];
    double s = b[T.in_idx[p]];
    if (OP == OP_BWD) s = s / (dp + tfs::kTiny);
    for (int e = e0; e < e1; ++e) s = s - T.val[e] * x[T.col[e]];
    x[p] = s * dp;
  }
}

template <int OP>
__global__ void __launch_bounds__(kLevelThreads)
k_level(int p0, int p1, Tri T, double *d, const double *__restrict__ b, double *x,
        const KState *st) {
  if (st && tfs::solve_done(st)) return;
  const int p = p0 + blockIdx.x * kLevelThreads + threadIdx.x;
  if (p < p1) sweep_row<OP>(T, d, b, x, p);
}

// A run of small levels [l0, l1) walked by one workgroup; writes of one level
// are visible to the whole workgroup after the barrier.
template <int OP>
__global__ void __launch_bounds__(kSmallThreads)
k_levels_small(int l0, int l1, const int *__restrict__ lev_ptr, Tri T, double *d,
               const double *__restrict__ b, double *x, const KState *st) {
  if (st && tfs::solve_done(st)) return;
  for (int l = l0; l < l1; ++l) {
    const int p1 = lev_ptr[l + 1];
    for (int p = lev_ptr[l] + threadIdx.x; p < p1; p += kSmallThreads) sweep_row<OP>(T, d, b, x, p);
    __syncthreads();
  }
}

// One persistent launch per sweep.  Workgroups take 256-position chunks in
// order from a counter (chunk c holds positions [256c, 256c+256)); inside a
// chunk the rows are handled level by level, and rows of level l start once
// every level below l is complete (per-level row counters).  A chunk only
// waits on positions before its own, all of which belong to chunks already
// taken by running workgroups, so the schedule cannot deadlock whatever the
// grid size.  Each thread loads its row's entries before it waits, so only
// the gathers of earlier results sit on the critical path.  Every wait is
// bounded: past kSpinLimit polls the kernel records an error and moves on,
// so it always terminates.  The last workgroup to leave resets the counters.
//   ctr[0] chunk counter, ctr[1] exit counter, ctr[2 + l] rows done in level l
constexpr int kChunk = 256;
constexpr int kPrefetch = 8;          // row entries held in registers
constexpr unsigned kSpinLimit = 1u << 22;

template <int OP>
__global__ void __launch_bounds__(kChunk)
k_sweep_persistent(int n, int nlev, const int *__restrict__ lev_ptr, Tri T, double *d,
                   const double *__restrict__ b, double *x, const KState *st,
                   unsigned int *ctr, int *err) {
  if (st && tfs::solve_done(st)) return;
  __shared__ int s_chunk, s_lmin, s_ready;
  unsigned int *cnt = ctr + 2;
  double *out = OP == OP_FACTOR ? d : x;
  const int nchunks = (n + kChunk - 1) / kChunk;
  if (threadIdx.x == 0) s_ready = 0;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_chunk = (int)atomicAdd(&ctr[0], 1u);
    __syncthreads();
    const int c = s_chunk;
    if (c >= nchunks) break;
    const int c0 = c * kChunk, c1 = min(n, c0 + kChunk);
    const int p = c0 + (int)threadIdx.x;
    // independent loads first: the row's entries and right-hand side
    int e0 = 0, e1 = 0;
    int pc[kPrefetch];
    double pv[kPrefetch];
    double rhs = 0.0, dp = 0.0;
    if (p < c1) {
      e0 = T.ptr[p];
      e1 = T.ptr[p + 1];
#pragma unroll
      for (int k = 0; k < kPrefetch; ++k) {
        if (e0 + k < e1) {
          pc[k] = T.col[e0 + k];
          pv[k] = T.val[e0 + k];
        }
      }
      if (OP == OP_FACTOR) {
        rhs = T.diag[p];
      } else {
        rhs = b[T.in_idx[p]];
        dp = d[p];
      }
    }
    if (threadIdx.x == 0) {
      int lo = 0, hi = nlev - 1;  // level of the chunk's first position
      while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (lev_ptr[mid] <= c0) lo = mid; else hi = mid - 1;
      }
      s_lmin = lo;
    }
    __syncthreads();
    for (int l = s_lmin; l < nlev && lev_ptr[l] < c1; ++l) {
      const int a = max(c0, lev_ptr[l]), z = min(c1, lev_ptr[l + 1]);
      if (threadIdx.x == 0) {
        int r = s_ready;
        unsigned int spins = 0;
        bool waited = false;
        while (r < l) {
          const unsigned int need = (unsigned int)(lev_ptr[r + 1] - lev_ptr[r]);
          if (__hip_atomic_load(&cnt[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= need) {
            ++r;
            continue;
          }
          waited = true;
          if (++spins > kSpinLimit) {
            *err = 1;
            r = l;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        if (waited || r != s_ready) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ready = r;
      }
      __syncthreads();
      if (p >= a && p < z) {
        // entries in the caller's order: the prefetched ones (static register
        // indices), then any beyond kPrefetch straight from memory
        const int len = e1 - e0;
        double s = rhs;
        if (OP == OP_BWD) s = s / (dp + tfs::kTiny);
#pragma unroll
        for (int k = 0; k < kPrefetch; ++k) {
          if (k < len) {
            if (OP == OP_FACTOR)
              s = s - out[pc[k]] * pv[k] * pv[k];
            else
              s = s - pv[k] * out[pc[k]];
          }
        }
        for (int e = e0 + kPrefetch; e < e1; ++e) {
          if (OP == OP_FACTOR)
            s = s - out[T.col[e]] * T.val[e] * T.val[e];
          else
            s = s - T.val[e] * out[T.col[e]];
        }
        out[p] = OP == OP_FACTOR ? 1.0 / s : s * dp;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(&cnt[l], (unsigned int)(z - a), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
  if (threadIdx.x == 0) {
    const unsigned int prev = __hip_atomic_fetch_add(&ctr[1], 1u, __ATOMIC_ACQ_REL,
                                                     __HIP_MEMORY_SCOPE_AGENT);
    if (prev == gridDim.x - 1) {  // everyone else has left: reset for the next launch
      for (int l = 0; l < nlev; ++l) cnt[l] = 0u;
      ctr[0] = 0u;
      ctr[1] = 0u;
    }
  }
}

// ---------------------------------------------------------------------------
// Host side.

struct Lib {
  std::mutex mu;
  bool ready = false;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};
Lib g_lib;

// Optional HIP-event timing of one kernel class on the library stream: an
// event pair brackets every launch of that class; pairs are harvested at the
// next stream synchronisation.
struct Timing {
  bool on = false;
  int kclass = -1;
  std::vector<hipEvent_t> ev;
  size_t used = 0;
  double ms = 0.0;
  long long launches = 0;
};
Timing g_tm;

inline bool tm_active(int kc) {
  return g_tm.on && g_tm.kclass == kc && g_tm.used + 2 <= g_tm.ev.size();
}
inline void tm_begin(int kc) {
  if (tm_active(kc)) (void)hipEventRecord(g_tm.ev[g_tm.used], g_lib.stream);
}
inline void tm_end(int kc) {
  if (tm_active(kc)) {
    (void)hipEventRecord(g_tm.ev[g_tm.used + 1], g_lib.stream);
    g_tm.used += 2;
  }
}
void tm_harvest() {  // stream idle
  for (size_t k = 0; k + 1 < g_tm.used; k += 2) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, g_tm.ev[k], g_tm.ev[k + 1]) == hipSuccess) {
      g_tm.ms += t;
      g_tm.launches++;
    }
  }
  g_tm.used = 0;
}

#define TFS_HIP(call)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) {                                                              \
      std::fprintf(stderr, "tfs: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(e_), \
                   __FILE__, __LINE__);                                                  \
      return e_ == hipErrorOutOfMemory ? TFS_ERR_NOMEM : TFS_ERR_DEVICE;                 \
    }                                                                                    \
  } while (0)

#define TFS_TRY(expr)        \
  do {                       \
    int rc_ = (expr);        \
    if (rc_ < 0) return rc_; \
  } while (0)

struct DevSweep {
  int nlev = 0;
  std::vector<int> lev_ptr;           // host copy
  std::vector<tfs::Segment> segs;
  int *lev = nullptr, *ptr = nullptr, *col = nullptr, *src = nullptr, *diag_src = nullptr;
  int *of_orig = nullptr;
  double *val = nullptr;              // triangle values
  long long ntri = 0;
  unsigned int *ctr = nullptr;        // persistent-sweep counters (2 + nlev)
};

}  // namespace

struct tfs_matrix {
  int n = 0;
  long long nnz = 0;
  // S layout
  int *s_row = nullptr, *s_col = nullptr, *s_src = nullptr, *s_diag_src = nullptr;
  int *s_orig = nullptr, *s_of_orig = nullptr, *rblk = nullptr;
  double *s_val = nullptr, *s_diag = nullptr;
  int nblk = 0;
  // L and U layouts
  DevSweep L, U;
  int *l_to_s = nullptr, *u_to_l = nullptr, *s_to_u = nullptr;
  double *l_diag = nullptr;    // a_ii in L order
  double *dL = nullptr, *dU = nullptr;  // factor in L and U order
  // caller-order values and the state of the derived copies
  double *val = nullptr;
  bool values_ready = false;   // S/L/U copies match val
  int factor_prec = -1;        // prec the factor was built for, -1 = stale
  // solve workspace
  double *vec = nullptr;
  int n_vec = 0;
  double *stage_x = nullptr, *stage_r = nullptr;  // host-memory solves
  KState *d_st = nullptr;
  KState *h_st = nullptr;      // pinned
  double *partials = nullptr;
  int *sweep_err = nullptr;    // set by a persistent sweep whose wait ran out
  bool persistent = true;      // one launch per sweep (else one per level)
  int grid_persistent = 0;
  long long bytes = 0;
  int bandwidth = 0;
  std::vector<void *> allocs;
};

namespace {

int dev_alloc(tfs_matrix *A, void **p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  TFS_HIP(hipMalloc(p, bytes));
  A->allocs.push_back(*p);
  A->bytes += (long long)bytes;
  return 0;
}

template <typename T>
int upload(tfs_matrix *A, T **dst, const std::vector<T> &src) {
  TFS_TRY(dev_alloc(A, (void **)dst, sizeof(T) * src.size()));
  if (!src.empty())
    TFS_HIP(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
  return 0;
}

int ensure_init() {
  std::lock_guard<std::mutex> lk(g_lib.mu);
  if (g_lib.ready) return 0;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return TFS_ERR_NO_DEVICE;
  if (g_lib.device >= ndev) return TFS_ERR_INVALID;
  TFS_HIP(hipSetDevice(g_lib.device));
  TFS_HIP(hipStreamCreateWithFlags(&g_lib.stream, hipStreamNonBlocking));
  TFS_HIP(hipEventCreate(&g_lib.ev0));
  TFS_HIP(hipEventCreate(&g_lib.ev1));
  g_lib.ready = true;
  return 0;
}

inline int grid_for(long long items, int per_block, int cap = kMaxGrid) {
  long long g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}
inline int vgrid(int n) { return grid_for(n, kRedThreads * 4); }

void gather(int n, const double *in, const int *idx, double *out, const KState *st = nullptr) {
  hipLaunchKernelGGL(k_gather, dim3(vgrid(n)), dim3(kRedThreads), 0, g_lib.stream, n, in, idx,
                     out, st);
}
void scatter(int n, const double *in, const int *idx, double *out) {
  hipLaunchKernelGGL(k_scatter, dim3(vgrid(n)), dim3(kRedThreads), 0, g_lib.stream, n, in, idx,
                     out);
}
void finish(int nparts, const double *partials, int epi, KState *st) {
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(kRedThreads), 0, g_lib.stream, nparts, partials, epi,
                     st);
}

template <int OP>
void launch_sweep(const tfs_matrix *A, const DevSweep &S, const Tri &T, double *d,
                  const double *b, double *x, const KState *st) {
  const int kc = OP == OP_FACTOR ? TFS_KC_FACTOR : TFS_KC_SWEEP;
  if (A->persistent) {
    tm_begin(kc);
    hipLaunchKernelGGL(k_sweep_persistent<OP>, dim3(A->grid_persistent), dim3(kChunk), 0,
                       g_lib.stream, A->n, S.nlev, S.lev, T, d, b, x, st, S.ctr, A->sweep_err);
    tm_end(kc);
    return;
  }
  for (const tfs::Segment &g : S.segs) {
    tm_begin(kc);
    if (g.small) {
      hipLaunchKernelGGL(k_levels_small<OP>, dim3(1), dim3(kSmallThreads), 0, g_lib.stream, g.l0,
                         g.l1, S.lev, T, d, b, x, st);
    } else {
      const int p0 = S.lev_ptr[g.l0], p1 = S.lev_ptr[g.l1];
      hipLaunchKernelGGL(k_level<OP>, dim3((p1 - p0 + kLevelThreads - 1) / kLevelThreads),
                         dim3(kLevelThreads), 0, g_lib.stream, p0, p1, T, d, b, x, st);
    }
    tm_end(kc);
  }
}

// Copies of the caller's values into the S, L and U layouts (part of the
// work the reference does in Prec_Form each call: the matrix changed).
void prepare_values(tfs_matrix *A) {
  if (A->values_ready) return;
  const int n = A->n;
  gather((int)A->nnz, A->val, A->s_src, A->s_val);
  gather(n, A->val, A->s_diag_src, A->s_diag);
  if (A->L.ntri) gather((int)A->L.ntri, A->val, A->L.src, A->L.val);
  if (A->U.ntri) gather((int)A->U.ntri, A->val, A->U.src, A->U.val);
  gather(n, A->val, A->L.diag_src, A->l_diag);
  A->values_ready = true;
}

int prec_form(tfs_matrix *A, int prec) {
  if (prec < 0 || prec > 2) prec = 0;  // the reference treats any other value as none
  prepare_values(A);
  if (A->factor_prec == prec) return 0;
  if (A->n > 0 && prec == 2) {
    const Tri T{A->L.ptr, A->L.col, A->L.val, A->l_diag, nullptr};
    launch_sweep<OP_FACTOR>(A, A->L, T, A->dL, nullptr, nullptr, nullptr);
    gather(A->n, A->dL, A->u_to_l, A->dU);  // the factor in U order
    TFS_HIP(hipGetLastError());
  }
  A->factor_prec = prec;
  return 0;
}

// zU := M^-1 b for prec 2, b in S order, result in U order (yL scratch).
void dic_apply(const tfs_matrix *A, const double *bS, double *yL, double *zU, const KState *st) {
  const Tri TL{A->L.ptr, A->L.col, A->L.val, nullptr, A->l_to_s};
  launch_sweep<OP_FWD>(A, A->L, TL, A->dL, bS, yL, st);
  const Tri TU{A->U.ptr, A->U.col, A->U.val, nullptr, A->u_to_l};
  launch_sweep<OP_BWD>(A, A->U, TU, A->dU, yL, zU, st);
}

// xS := M^-1 bS, both in S order (CGS and the building block).
void prec_apply_S(const tfs_matrix *A, int prec, const double *bS, double *xS, double *yL,
                  double *zU, const KState *st) {
  if (prec == 2) {
    dic_apply(A, bS, yL, zU, st);
    gather(A->n, zU, A->s_to_u, xS, st);
  } else {
    hipLaunchKernelGGL(k_diag_apply, dim3(vgrid(A->n)), dim3(kRedThreads), 0, g_lib.stream, A->n,
                       bS, A->s_diag, xS, prec, st);
  }
}

template <int MODE>
void launch_spmv(const tfs_matrix *A, const double *x, double *y, const double *w, double *xs,
                 double *rs, const KState *st) {
  tm_begin(TFS_KC_SPMV);
  hipLaunchKernelGGL(k_spmv<MODE>, dim3(grid_for(A->nblk, kSpmvWaves)), dim3(kSpmvThreads), 0,
                     g_lib.stream, A->nblk, A->rblk, A->s_row, A->s_col, A->s_val, x, y, w, xs,
                     rs, st, A->partials);
  tm_end(TFS_KC_SPMV);
}
inline int spmv_parts(const tfs_matrix *A) { return grid_for(A->nblk, kSpmvWaves); }

int ensure_vectors(tfs_matrix *A, int count) {
  if (A->n_vec >= count) return 0;
  if (A->vec) {
    TFS_HIP(hipFree(A->vec));
    A->allocs.erase(std::find(A->allocs.begin(), A->allocs.end(), (void *)A->vec));
    A->bytes -= (long long)A->n_vec * A->n * (long long)sizeof(double);
    A->vec = nullptr;
  }
  TFS_TRY(dev_alloc(A, (void **)&A->vec, sizeof(double) * (size_t)A->n * count));
  A->n_vec = count;
  return 0;
}

// Workspace vectors (S order unless noted).
struct Vecs {
  double *x, *r, *z, *p, *q, *yL, *zU;                 // CG
  double *p1, *p2, *q1, *r2, *u1, *u2, *v2, *w;        // CGS (aliases above where free)
};
Vecs vecs(tfs_matrix *A) {
  const size_t n = (size_t)A->n;
  double *v = A->vec;
  Vecs V;
  V.x = v; V.r = v + n; V.yL = v + 2 * n; V.zU = v + 3 * n;
  V.z = v + 4 * n; V.p = v + 5 * n; V.q = v + 6 * n;
  V.p1 = v + 4 * n; V.p2 = v + 5 * n; V.q1 = v + 6 * n; V.r2 = v + 7 * n;
  V.u1 = v + 8 * n; V.u2 = v + 9 * n; V.v2 = v + 10 * n; V.w = v + 11 * n;
  return V;
}

// One CG iteration (Cg.f90:97-178); Bicg takes the same path (its shadow
// sequence equals the primary one bit for bit, Bicg.f90:91-175).
void cg_iteration(tfs_matrix *A, int prec, const Vecs &V) {
  KState *st = A->d_st;
  const int n = A->n, g = vgrid(n);
  if (prec == 2) dic_apply(A, V.r, V.yL, V.zU, st);
  const int mode = prec == 2 ? 2 : prec == 1 ? 1 : 0;
  double *z = prec == 0 ? V.r : V.z;  // prec 0: z is a copy of r (Prec_Solve.f90:72-75)
  hipLaunchKernelGGL(k_rho, dim3(g), dim3(kRedThreads), 0, g_lib.stream, n, V.r, V.z, A->s_diag,
                     V.zU, A->s_to_u, mode, st, A->partials);
  finish(g, A->partials, EPI_RHO, st);
  hipLaunchKernelGGL(k_pupdate, dim3(g), dim3(kRedThreads), 0, g_lib.stream, n, z, V.p, st);
  launch_spmv<SPMV_DOT>(A, V.p, V.q, V.p, nullptr, nullptr, st);
  finish(spmv_parts(A), A->partials, EPI_ALPHA, st);
  hipLaunchKernelGGL(k_axpy_norm, dim3(g), dim3(kRedThreads), 0, g_lib.stream, n, V.x, V.r, V.p,
                     V.q, st, A->partials);
  finish(g, A->partials, EPI_NORM, st);
}

// One CGS iteration (Cgs.f90:104-226).
void cgs_iteration(tfs_matrix *A, int prec, const Vecs &V) {
  KState *st = A->d_st;
  const int n = A->n, g = vgrid(n);
  hipLaunchKernelGGL(k_rho, dim3(g), dim3(kRedThreads), 0, g_lib.stream, n, V.r, V.r2, nullptr,
                     nullptr, nullptr, 3, st, A->partials);  // rho = (r1, r2)
  finish(g, A->partials, EPI_RHO, st);
  hipLaunchKernelGGL(k_cgs_u, dim3(g), dim3(kRedThreads), 0, g_lib.stream, n, V.r, V.q1, V.u1,
                     V.u2, st);
  double *pp2 = V.u2;  // prec 0: p2 is a copy of u2
  if (prec != 0) {
    prec_apply_S(A, prec, V.u2, V.p2, V.yL, V.zU, st);
    pp2 = V.p2;
  }
  launch_spmv<SPMV_DOT>(A, pp2, V.v2, V.r2, nullptr, nullptr, st);  // v2 = A p2, (r2, v2)
  finish(spmv_parts(A), A->partials, EPI_ALPHA, st);
  hipLaunchKernelGGL(k_cgs_q, dim3(g), dim3(kRedThreads), 0, g_lib.stream, n, V.u1, V.v2, V.q1,
                     V.w, st);
  double *pp1 = V.w;
  if (prec != 0) {
    prec_apply_S(A, prec, V.w, V.p1, V.yL, V.zU, st);
    pp1 = V.p1;
  }
  // q2 = A p1 into the v2 slot (v2 is dead once q1 and w are formed)
  launch_spmv<SPMV_CGS_TAIL>(A, pp1, V.v2, nullptr, V.x, V.r, st);
  finish(spmv_parts(A), A->partials, EPI_NORM, st);
}

// A persistent sweep that gave up waiting leaves a flag; report it (the
// results of that call are not trustworthy) and clear it.
int check_sweep(tfs_matrix *A) {
  int e = 0;
  TFS_HIP(hipMemcpy(&e, A->sweep_err, sizeof(int), hipMemcpyDeviceToHost));
  if (e) {
    TFS_HIP(hipMemset(A->sweep_err, 0, sizeof(int)));
    std::fprintf(stderr, "tfs: a preconditioner sweep timed out waiting for its dependencies\n");
    return TFS_ERR_DEVICE;
  }
  return 0;
}

int read_state(tfs_matrix *A) {
  TFS_HIP(hipMemcpyAsync(A->h_st, A->d_st, sizeof(KState), hipMemcpyDeviceToHost, g_lib.stream));
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  if (g_tm.on) tm_harvest();
  return 0;
}

int solve_device(tfs_matrix *A, int method, int prec, double *x, double *r1, int *niter,
                 double tol, double *ini_res, double *fin_res) {
  if (!A || !niter || !ini_res || !fin_res) return TFS_ERR_INVALID;
  if (method < TFS_CG || method > TFS_CGS) return TFS_ERR_INVALID;
  if (prec < 0 || prec > 2) prec = 0;
  const int n = A->n;
  const int cap = *niter;
  if (n == 0) {  // ||b|| = 0
    *niter = 0;
    *fin_res = 0.0;
    return TFS_OK;
  }
  if (!x || !r1) return TFS_ERR_INVALID;
  TFS_TRY(ensure_vectors(A, method == TFS_CGS ? 12 : 7));
  TFS_TRY(prec_form(A, prec));  // Cg.f90:42 (kept while the values are unchanged)
  const Vecs V = vecs(A);

  KState *h = A->h_st;
  std::memset(h, 0, sizeof(KState));
  h->tol = tol;
  h->cap = cap;
  h->ini_res = *ini_res;
  TFS_HIP(hipMemcpyAsync(A->d_st, h, sizeof(KState), hipMemcpyHostToDevice, g_lib.stream));
  KState *st = A->d_st;
  const int g = vgrid(n);
  // caller order -> S order
  gather(n, x, A->s_orig, V.x);
  gather(n, r1, A->s_orig, V.r);
  // ||b|| and its early exit; r = b - A x; ||r|| and its early exit
  hipLaunchKernelGGL(k_sumsq, dim3(g), dim3(kRedThreads), 0, g_lib.stream, n, V.r, (int)EPI_BNRM,
                     st, A->partials);
  finish(g, A->partials, EPI_BNRM, st);
  launch_spmv<SPMV_RESID>(A, V.x, V.r, nullptr, nullptr, nullptr, st);
  if (method == TFS_CGS) gather(n, V.r, nullptr, V.r2, st);  // r2 = r1 (Cgs.f90:75-77)
  hipLaunchKernelGGL(k_sumsq, dim3(g), dim3(kRedThreads), 0, g_lib.stream, n, V.r,
                     (int)EPI_INIRES, st, A->partials);
  finish(g, A->partials, EPI_INIRES, st);
  TFS_HIP(hipGetLastError());
  TFS_TRY(read_state(A));

  if (!h->done && cap < 1) {  // zero-trip DO loop: iter = 1 (Cg.f90:97,181)
    h->done = 1;
    h->niter_out = 1;
    h->fin_res = h->err;
    h->status = tfs::kStNotConverged;
  }
  int queued = 0, chunk = 1, it_prev = 0;
  double err_prev = h->err;
  while (!h->done) {
    for (int k = 0; k < chunk; ++k) {
      if (method == TFS_CGS)
        cgs_iteration(A, prec, V);
      else
        cg_iteration(A, prec, V);
    }
    queued += chunk;
    TFS_HIP(hipGetLastError());
    TFS_TRY(read_state(A));
    if (h->done) break;
    if (queued > cap + 64) return TFS_ERR_DEVICE;  // the norm epilogue stops every solve at cap
    // Next chunk: the iterations still needed at the observed contraction
    // rate, so that few are queued past convergence.
    const int dit = h->iter - it_prev;
    int next = 16;
    if (dit > 0 && h->err > 0 && err_prev > 0 && h->err < err_prev) {
      const double rate = std::pow(h->err / err_prev, 1.0 / dit);
      const double need = std::log(h->tol / h->err) / std::log(rate);
      next = need < 1.0 ? 1 : (int)(0.9 * need);
    }
    chunk = std::max(1, std::min(32, next));
    err_prev = h->err;
    it_prev = h->iter;
  }
  // S order -> caller order (x and the final residual in r1)
  scatter(n, V.x, A->s_orig, x);
  scatter(n, V.r, A->s_orig, r1);
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  TFS_TRY(check_sweep(A));
  *niter = h->niter_out;
  *fin_res = h->fin_res;
  *ini_res = h->ini_res;  // still the caller's value after the ||b|| < tol exit
  if (h->status == tfs::kStBreakdown) return TFS_ERR_BREAKDOWN;
  return h->status == tfs::kStNotConverged ? TFS_NOT_CONVERGED : TFS_OK;
}

int upload_sweep(tfs_matrix *A, const tfs::SweepLayout &S, DevSweep &D) {
  D.nlev = S.nlev;
  D.lev_ptr = S.lev_ptr;
  D.segs = S.segs;
  D.ntri = (long long)S.col.size();
  TFS_TRY(upload(A, &D.lev, S.lev_ptr));
  TFS_TRY(upload(A, &D.ptr, S.ptr));
  TFS_TRY(upload(A, &D.col, S.col));
  TFS_TRY(upload(A, &D.src, S.src));
  TFS_TRY(upload(A, &D.diag_src, S.diag_src));
  TFS_TRY(upload(A, &D.of_orig, S.of_orig));
  TFS_TRY(dev_alloc(A, (void **)&D.val, sizeof(double) * (size_t)std::max<long long>(D.ntri, 1)));
  TFS_TRY(dev_alloc(A, (void **)&D.ctr, sizeof(unsigned int) * (size_t)(2 + D.nlev)));
  TFS_HIP(hipMemset(D.ctr, 0, sizeof(unsigned int) * (size_t)(2 + D.nlev)));
  return 0;
}

int create_device(tfs_matrix *A, const tfs::Layout &Ly) {
  const int n = A->n;
  TFS_TRY(upload(A, &A->s_row, Ly.s_row));
  TFS_TRY(upload(A, &A->s_col, Ly.s_col));
  TFS_TRY(upload(A, &A->s_src, Ly.s_src));
  TFS_TRY(upload(A, &A->s_diag_src, Ly.s_diag_src));
  TFS_TRY(upload(A, &A->s_orig, Ly.s_orig));
  TFS_TRY(upload(A, &A->s_of_orig, Ly.s_of_orig));
  TFS_TRY(upload(A, &A->rblk, Ly.s_rblk));
  A->nblk = (int)Ly.s_rblk.size() - 1;
  TFS_TRY(upload_sweep(A, Ly.L, A->L));
  TFS_TRY(upload_sweep(A, Ly.U, A->U));
  TFS_TRY(upload(A, &A->l_to_s, Ly.l_to_s));
  TFS_TRY(upload(A, &A->u_to_l, Ly.u_to_l));
  TFS_TRY(upload(A, &A->s_to_u, Ly.s_to_u));
  const size_t nb = sizeof(double) * (size_t)std::max(n, 1);
  TFS_TRY(dev_alloc(A, (void **)&A->s_val, sizeof(double) * (size_t)std::max(A->nnz, 1LL)));
  TFS_TRY(dev_alloc(A, (void **)&A->val, sizeof(double) * (size_t)std::max(A->nnz, 1LL)));
  TFS_TRY(dev_alloc(A, (void **)&A->s_diag, nb));
  TFS_TRY(dev_alloc(A, (void **)&A->l_diag, nb));
  TFS_TRY(dev_alloc(A, (void **)&A->dL, nb));
  TFS_TRY(dev_alloc(A, (void **)&A->dU, nb));
  TFS_TRY(dev_alloc(A, (void **)&A->stage_x, nb));
  TFS_TRY(dev_alloc(A, (void **)&A->stage_r, nb));
  TFS_TRY(dev_alloc(A, (void **)&A->d_st, sizeof(KState)));
  TFS_TRY(dev_alloc(A, (void **)&A->partials, sizeof(double) * kMaxGrid));
  TFS_TRY(dev_alloc(A, (void **)&A->sweep_err, sizeof(int)));
  TFS_HIP(hipMemset(A->sweep_err, 0, sizeof(int)));
  {
    int dev = 0, n_cu = 0, per_cu = 0;
    TFS_HIP(hipGetDevice(&dev));
    TFS_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    TFS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void *>(k_sweep_persistent<OP_FWD>), kChunk, 0));
    const int chunks = (n + kChunk - 1) / kChunk;
    A->grid_persistent = std::max(1, std::min(chunks, n_cu * std::max(1, std::min(per_cu, 8))));
  }
  if (const char *e = std::getenv("TFS_SWEEP")) A->persistent = std::strcmp(e, "levels") != 0;
  TFS_HIP(hipHostMalloc((void **)&A->h_st, sizeof(KState), hipHostMallocDefault));
  TFS_HIP(hipMemset(A->d_st, 0, sizeof(KState)));
  TFS_HIP(hipMemset(A->val, 0, sizeof(double) * (size_t)std::max(A->nnz, 1LL)));
  A->bandwidth = Ly.bandwidth;
  return 0;
}

}  // namespace

// ============================================================================
// C ABI

extern "C" const char *tfs_status_string(int s) {
  switch (s) {
    case TFS_OK: return "ok";
    case TFS_NOT_CONVERGED: return "not converged (niter = cap+1)";
    case TFS_ERR_INVALID: return "invalid argument or CRS structure";
    case TFS_ERR_NOMEM: return "device out of memory";
    case TFS_ERR_BREAKDOWN: return "breakdown (zero, Inf or NaN in a Krylov scalar)";
    case TFS_ERR_DEVICE: return "HIP runtime error";
    case TFS_ERR_NO_DEVICE: return "no GPU visible";
    default: return "unknown status";
  }
}

extern "C" int tfs_init(int device) {
  {
    std::lock_guard<std::mutex> lk(g_lib.mu);
    if (g_lib.ready) return device == g_lib.device ? TFS_OK : TFS_ERR_INVALID;
    g_lib.device = device;
  }
  return ensure_init();
}

extern "C" int tfs_matrix_create(int n, long long nnz, const int *row, const int *col,
                                 const int *dia, int base, tfs_matrix **out) {
  if (!out) return TFS_ERR_INVALID;
  *out = nullptr;
  if (n < 0 || nnz < n || nnz > 0x7ffffffeLL || (base != 0 && base != 1)) return TFS_ERR_INVALID;
  if (n > 0 && (!row || !col || !dia)) return TFS_ERR_INVALID;

  // Validate and rebase the structure (Matrix_Mod.f90:15-34 layout) before
  // touching the device, so a bad structure is rejected without a GPU.
  std::vector<int> hr(n + 1), hc((size_t)nnz), hd(n);
  for (int i = 0; i <= n; ++i) hr[i] = row[i] - base;
  if (hr[0] != 0 || hr[n] != nnz) return TFS_ERR_INVALID;
  for (int i = 0; i < n; ++i)
    if (hr[i + 1] < hr[i] + 1) return TFS_ERR_INVALID;
  for (long long k = 0; k < nnz; ++k) {
    hc[(size_t)k] = col[k] - base;
    if (hc[(size_t)k] < 0 || hc[(size_t)k] >= n) return TFS_ERR_INVALID;
  }
  for (int i = 0; i < n; ++i) {
    hd[i] = dia[i] - base;
    if (hd[i] < hr[i] || hd[i] >= hr[i + 1] || hc[(size_t)hd[i]] != i) return TFS_ERR_INVALID;
    for (int j = hr[i] + 1; j < hr[i + 1]; ++j)
      if (hc[(size_t)j] <= hc[(size_t)j - 1]) return TFS_ERR_INVALID;  // ascending, distinct
  }
  TFS_TRY(ensure_init());

  tfs::Layout Ly;
  tfs::build_layout(n, hr.data(), hc.data(), hd.data(), kSpmvRows, kSpmvCap, kSmallLevel, Ly);
  tfs_matrix *A = new tfs_matrix;
  A->n = n;
  A->nnz = nnz;
  const int rc = create_device(A, Ly);
  if (rc < 0) {
    tfs_matrix_destroy(A);
    return rc;
  }
  *out = A;
  return TFS_OK;
}

extern "C" int tfs_matrix_destroy(tfs_matrix *A) {
  if (!A) return TFS_OK;
  if (g_lib.stream) (void)hipStreamSynchronize(g_lib.stream);
  for (void *p : A->allocs) (void)hipFree(p);
  if (A->h_st) (void)hipHostFree(A->h_st);
  delete A;
  return TFS_OK;
}

extern "C" int tfs_matrix_set_values(tfs_matrix *A, const double *val) {
  if (!A || (A->nnz > 0 && !val)) return TFS_ERR_INVALID;
  TFS_HIP(hipMemcpyAsync(A->val, val, sizeof(double) * (size_t)A->nnz, hipMemcpyHostToDevice,
                         g_lib.stream));
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  A->values_ready = false;
  A->factor_prec = -1;
  return TFS_OK;
}

extern "C" int tfs_matrix_set_values_device(tfs_matrix *A, const double *d_val) {
  if (!A || (A->nnz > 0 && !d_val)) return TFS_ERR_INVALID;
  TFS_HIP(hipMemcpyAsync(A->val, d_val, sizeof(double) * (size_t)A->nnz, hipMemcpyDeviceToDevice,
                         g_lib.stream));
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  A->values_ready = false;
  A->factor_prec = -1;
  return TFS_OK;
}

extern "C" int tfs_matrix_values_changed(tfs_matrix *A) {
  if (!A) return TFS_ERR_INVALID;
  A->values_ready = false;
  A->factor_prec = -1;
  return TFS_OK;
}

extern "C" double *tfs_matrix_values_device(tfs_matrix *A) { return A ? A->val : nullptr; }

extern "C" int tfs_matrix_get_info(const tfs_matrix *A, tfs_matrix_info *info) {
  if (!A || !info) return TFS_ERR_INVALID;
  info->n = A->n;
  info->nnz = A->nnz;
  info->levels_lower = A->L.nlev;
  info->levels_upper = A->U.nlev;
  info->row_blocks = A->nblk;
  info->device_bytes = A->bytes;
  return TFS_OK;
}

extern "C" int tfs_solve_device(tfs_matrix *A, int method, int prec, double *d_x, double *d_r1,
                                int *niter, double tol, double *ini_res, double *fin_res) {
  return solve_device(A, method, prec, d_x, d_r1, niter, tol, ini_res, fin_res);
}

extern "C" int tfs_solve(tfs_matrix *A, int method, int prec, double *x1, double *r1, int *niter,
                         double tol, double *ini_res, double *fin_res) {
  if (!A) return TFS_ERR_INVALID;
  if (A->n > 0 && (!x1 || !r1)) return TFS_ERR_INVALID;
  const size_t bytes = sizeof(double) * (size_t)A->n;
  if (A->n > 0) {
    TFS_HIP(hipMemcpyAsync(A->stage_x, x1, bytes, hipMemcpyHostToDevice, g_lib.stream));
    TFS_HIP(hipMemcpyAsync(A->stage_r, r1, bytes, hipMemcpyHostToDevice, g_lib.stream));
  }
  const int rc = solve_device(A, method, prec, A->stage_x, A->stage_r, niter, tol, ini_res, fin_res);
  if (rc < 0 && rc != TFS_ERR_BREAKDOWN) return rc;
  if (A->n > 0) {
    TFS_HIP(hipMemcpyAsync(x1, A->stage_x, bytes, hipMemcpyDeviceToHost, g_lib.stream));
    TFS_HIP(hipMemcpyAsync(r1, A->stage_r, bytes, hipMemcpyDeviceToHost, g_lib.stream));
    TFS_HIP(hipStreamSynchronize(g_lib.stream));
  }
  return rc;
}

// Building blocks on caller-order vectors: permuted into the S layout and back.
extern "C" int tfs_spmv_device(tfs_matrix *A, const double *d_x, double *d_y) {
  if (!A || (A->n > 0 && (!d_x || !d_y))) return TFS_ERR_INVALID;
  if (A->n == 0) return TFS_OK;
  TFS_TRY(ensure_vectors(A, 7));
  prepare_values(A);
  const Vecs V = vecs(A);
  gather(A->n, d_x, A->s_orig, V.x);
  launch_spmv<SPMV_PLAIN>(A, V.x, V.q, nullptr, nullptr, nullptr, nullptr);
  scatter(A->n, V.q, A->s_orig, d_y);
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  return TFS_OK;
}

extern "C" int tfs_residual_device(tfs_matrix *A, const double *d_x, double *d_r) {
  if (!A || (A->n > 0 && (!d_x || !d_r))) return TFS_ERR_INVALID;
  if (A->n == 0) return TFS_OK;
  TFS_TRY(ensure_vectors(A, 7));
  prepare_values(A);
  const Vecs V = vecs(A);
  gather(A->n, d_x, A->s_orig, V.x);
  gather(A->n, d_r, A->s_orig, V.r);
  launch_spmv<SPMV_RESID>(A, V.x, V.r, nullptr, nullptr, nullptr, nullptr);
  scatter(A->n, V.r, A->s_orig, d_r);
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  return TFS_OK;
}

extern "C" int tfs_prec_form(tfs_matrix *A, int prec) {
  if (!A) return TFS_ERR_INVALID;
  TFS_TRY(prec_form(A, prec));
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  TFS_TRY(check_sweep(A));
  return TFS_OK;
}

extern "C" int tfs_prec_solve_device(tfs_matrix *A, int prec, const double *d_b, double *d_x) {
  if (!A || (A->n > 0 && (!d_b || !d_x))) return TFS_ERR_INVALID;
  if (prec < 0 || prec > 2) prec = 0;
  if (A->factor_prec != prec) return TFS_ERR_INVALID;  // tfs_prec_form first
  if (A->n == 0) return TFS_OK;
  TFS_TRY(ensure_vectors(A, 7));
  const Vecs V = vecs(A);
  gather(A->n, d_b, A->s_orig, V.r);
  prec_apply_S(A, prec, V.r, V.z, V.yL, V.zU, nullptr);
  scatter(A->n, V.z, A->s_orig, d_x);
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  TFS_TRY(check_sweep(A));
  return TFS_OK;
}

extern "C" int tfs_prec_diag_device(tfs_matrix *A, double *d_out) {
  if (!A || (A->n > 0 && !d_out)) return TFS_ERR_INVALID;
  if (A->factor_prec < 0) return TFS_ERR_INVALID;
  if (A->n == 0) return TFS_OK;
  if (A->factor_prec == 2) {
    gather(A->n, A->dL, A->L.of_orig, d_out);
  } else if (A->factor_prec == 1) {
    gather(A->n, A->s_diag, A->s_of_orig, d_out);
  } else {
    hipLaunchKernelGGL(k_fill, dim3(vgrid(A->n)), dim3(kRedThreads), 0, g_lib.stream, A->n, d_out,
                       1.0);
  }
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  return TFS_OK;
}

extern "C" int tfs_spmv_time(tfs_matrix *A, const double *d_x, double *d_y, int reps, double *ms) {
  if (!A || !ms || reps < 1 || (A->n > 0 && (!d_x || !d_y))) return TFS_ERR_INVALID;
  if (A->n == 0) {
    *ms = 0.0;
    return TFS_OK;
  }
  prepare_values(A);
  // The vectors are taken as S-ordered: the kernel's traffic is the same.
  launch_spmv<SPMV_PLAIN>(A, d_x, d_y, nullptr, nullptr, nullptr, nullptr);  // warm
  TFS_HIP(hipEventRecord(g_lib.ev0, g_lib.stream));
  for (int k = 0; k < reps; ++k)
    launch_spmv<SPMV_PLAIN>(A, d_x, d_y, nullptr, nullptr, nullptr, nullptr);
  TFS_HIP(hipEventRecord(g_lib.ev1, g_lib.stream));
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipEventSynchronize(g_lib.ev1));
  float t = 0.f;
  TFS_HIP(hipEventElapsedTime(&t, g_lib.ev0, g_lib.ev1));
  *ms = (double)t / reps;
  return TFS_OK;
}

extern "C" int tfs_timing_start(int kclass) {
  if (kclass < TFS_KC_SPMV || kclass > TFS_KC_VECTOR) return TFS_ERR_INVALID;
  TFS_TRY(ensure_init());
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  if (g_tm.ev.empty()) {
    g_tm.ev.resize(2 * 4096);
    for (auto &e : g_tm.ev) TFS_HIP(hipEventCreate(&e));
  }
  g_tm.on = true;
  g_tm.kclass = kclass;
  g_tm.used = 0;
  g_tm.ms = 0.0;
  g_tm.launches = 0;
  return TFS_OK;
}

extern "C" int tfs_timing_stop(double *ms, long long *launches) {
  if (!g_tm.on) return TFS_ERR_INVALID;
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  tm_harvest();
  g_tm.on = false;
  if (ms) *ms = g_tm.ms;
  if (launches) *launches = g_tm.launches;
  return TFS_OK;
}
