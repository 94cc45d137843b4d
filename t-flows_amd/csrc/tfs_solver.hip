// tfs_solver.hip -- MI355X (gfx950) drop-in for T-FlowS's Krylov solvers.
//
// Kernels (all HBM-bound; none is a dense contraction, so no MFMA):
//   k_spmv          CSR SpMV, one wavefront per row block (<= 64 rows and
//                   <= kSpmvCap entries).  col/val are streamed with
//                   consecutive lanes on consecutive entries; the products
//                   val*x[col] are staged in LDS and each lane then sums its
//                   row left to right, which is the reference's order
//                   (Cg.f90:128-134), so y is bitwise the reference's.
//   k_*_level /     the diagonal incomplete Cholesky factor and the forward /
//   k_*_small       backward sweeps (Prec_Form.f90:32-40, Prec_Solve.f90:46-67),
//                   level-scheduled over the lower / upper dependency DAG:
//                   the rows of one level are independent.  Big levels get a
//                   launch each; runs of small levels are walked by one
//                   workgroup with a barrier between levels.
//   k_sumsq, k_rho, fused vector updates with deterministic grid reductions
//   k_axpy_norm...  (tfs_device.hpp); the scalars live on the device (KState)
//                   and each kernel returns at once once the solve is done,
//                   so iterations are queued ahead without host round trips.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "tfs_device.hpp"
#include "tfs_solver.h"

using tfs::KState;

namespace {

constexpr int kSpmvWaves = 4;            // waves per SpMV workgroup
constexpr int kSpmvThreads = kSpmvWaves * tfs::kWave;
constexpr int kSpmvCap = 512;            // entries per row block staged in LDS
constexpr int kRedThreads = 256;         // threads of the vector kernels
constexpr int kMaxGrid = 8192;           // upper bound of any reduction grid
constexpr int kLevelThreads = 256;       // threads per block, one level launch
constexpr int kSmallThreads = 1024;      // one workgroup walking small levels
constexpr int kSmallLevel = 2048;        // a level this small joins a walked run

enum Epi { EPI_NONE = 0, EPI_BNRM, EPI_INIRES, EPI_RHO, EPI_ALPHA, EPI_NORM };

// ---------------------------------------------------------------------------
// Solve-state transitions, run by thread 0 of the last workgroup of a
// reduction (Cg.f90:48-58,75-90,108-122,148-153,167-181).
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
    case EPI_RHO: {
      st->rho = tot;
      if (st->iter + 1 > 1) st->beta = tot / st->rho_old;
      break;
    }
    case EPI_ALPHA: {
      st->pq = tot;
      st->alpha = st->rho / tot;
      break;
    }
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
__device__ __forceinline__ void reduce_and_finish(double part, int epi, KState *st,
                                                  double *partials, unsigned int *ctr) {
  __shared__ double lds[NT / tfs::kWave];
  const double bs = tfs::block_sum<NT>(part, lds);
  double tot;
  if (tfs::grid_sum_last<NT>(bs, partials, ctr, &tot, lds) && threadIdx.x == 0)
    epilogue(epi, tot, st);
}

// ---------------------------------------------------------------------------
// Vector kernels (grid-stride; every one is a single pass over its vectors).

// sum of squares -> ||v|| epilogue (bnrm2 or the initial residual)
__global__ void __launch_bounds__(kRedThreads)
k_sumsq(int n, const double *__restrict__ v, int epi, KState *st, double *partials) {
  if (epi != EPI_BNRM && tfs::solve_done(st)) return;
  double s = 0.0;
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads)
    s += v[i] * v[i];
  reduce_and_finish<kRedThreads>(s, epi, st, partials, &st->ctr[0]);
}

// rho = (r, z); with jacobi != 0 also z = r / d first (Prec_Solve.f90:35-38).
__global__ void __launch_bounds__(kRedThreads)
k_rho(int n, const double *__restrict__ r, double *__restrict__ z, const double *__restrict__ d,
      int jacobi, KState *st, double *partials) {
  if (tfs::solve_done(st)) return;
  double s = 0.0;
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads) {
    double zi;
    if (jacobi) {
      zi = r[i] / d[i];
      z[i] = zi;
    } else {
      zi = z[i];
    }
    s += r[i] * zi;
  }
  reduce_and_finish<kRedThreads>(s, EPI_RHO, st, partials, &st->ctr[1]);
}

// p = z (first iteration) or p = z + beta p  (Cg.f90:114-123)
__global__ void __launch_bounds__(kRedThreads)
k_pupdate(int n, const double *__restrict__ z, double *__restrict__ p, const KState *st) {
  if (tfs::solve_done(st)) return;
  const bool first = st->iter + 1 == 1;
  const double beta = st->beta;
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads)
    p[i] = first ? z[i] : z[i] + beta * p[i];
}

// x += alpha p; r -= alpha q; ||r||  (Cg.f90:159-174)
__global__ void __launch_bounds__(kRedThreads)
k_axpy_norm(int n, double *__restrict__ x, double *__restrict__ r, const double *__restrict__ p,
            const double *__restrict__ q, KState *st, double *partials) {
  if (tfs::solve_done(st)) return;
  const double alpha = st->alpha;
  double s = 0.0;
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads) {
    x[i] = x[i] + alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    s += ri * ri;
  }
  reduce_and_finish<kRedThreads>(s, EPI_NORM, st, partials, &st->ctr[3]);
}

// CGS: u1 = r1 + beta q1; u2 = u1 + beta (q1 + beta u2)  (Cgs.f90:115-126)
__global__ void __launch_bounds__(kRedThreads)
k_cgs_u(int n, const double *__restrict__ r1, const double *__restrict__ q1,
        double *__restrict__ u1, double *__restrict__ u2, const KState *st) {
  if (tfs::solve_done(st)) return;
  const bool first = st->iter + 1 == 1;
  const double beta = st->beta;
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads) {
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
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads) {
    const double q = u1[i] - alpha * v2[i];
    q1[i] = q;
    w[i] = u1[i] + q;
  }
}

__global__ void __launch_bounds__(kRedThreads)
k_jacobi(int n, const double *__restrict__ b, const double *__restrict__ d, double *__restrict__ x,
         const KState *st) {
  if (st && tfs::solve_done(st)) return;
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads)
    x[i] = b[i] / d[i];
}

__global__ void __launch_bounds__(kRedThreads)
k_copy(int n, const double *__restrict__ b, double *__restrict__ x, const KState *st) {
  if (st && tfs::solve_done(st)) return;
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads)
    x[i] = b[i];
}

// Prec_Form.f90:24-27 (prec 1) and :45-48 (prec 0)
__global__ void __launch_bounds__(kRedThreads)
k_factor_diag(int n, const double *__restrict__ val, const int *__restrict__ dia,
              double *__restrict__ d, int prec) {
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads)
    d[i] = prec == 1 ? val[dia[i]] : 1.0;
}

// ---------------------------------------------------------------------------
// SpMV.  Modes: plain y = A x; residual y = y - A x (term by term,
// Residual.f90:22-27); dot: y = A x and sum w.y -> alpha (Cg.f90:148-153,
// Cgs.f90:157-162); cgs tail: q2 = A p1, x += alpha p1, r1 -= alpha q2 and
// ||r1|| (Cgs.f90:182-221, all row-local once alpha is known).
enum SpmvMode { SPMV_PLAIN = 0, SPMV_RESID, SPMV_DOT, SPMV_CGS_TAIL };

template <int MODE>
__global__ void __launch_bounds__(kSpmvThreads)
k_spmv(int nblk, const int *__restrict__ rblk, const int *__restrict__ row,
       const int *__restrict__ col, const double *__restrict__ val, const double *__restrict__ x,
       double *__restrict__ y, const double *__restrict__ w, double *__restrict__ xs,
       double *__restrict__ rs, KState *st, double *partials) {
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
  if (MODE == SPMV_DOT)
    reduce_and_finish<kSpmvThreads>(acc_red, EPI_ALPHA, st, partials, &st->ctr[2]);
  if (MODE == SPMV_CGS_TAIL)
    reduce_and_finish<kSpmvThreads>(acc_red, EPI_NORM, st, partials, &st->ctr[3]);
}

// ---------------------------------------------------------------------------
// Level-scheduled DIC: factor, forward and backward sweeps.

struct RowsCRS {
  const int *__restrict__ row;
  const int *__restrict__ col;
  const int *__restrict__ dia;
  const double *__restrict__ val;
};

// d_i = 1 / (a_ii - sum_{j<i} d_j a_ij a_ij)   (Prec_Form.f90:32-40)
__device__ __forceinline__ void factor_row(const RowsCRS &A, double *d, int i) {
  double s = A.val[A.dia[i]];
  for (int j = A.row[i]; j < A.dia[i]; ++j) s = s - d[A.col[j]] * A.val[j] * A.val[j];
  d[i] = 1.0 / s;
}
// x_i = (b_i - sum_{j<i} a_ij x_j) d_i   (Prec_Solve.f90:46-53)
__device__ __forceinline__ void fwd_row(const RowsCRS &A, const double *d, const double *b,
                                        double *x, int i) {
  double s = b[i];
  for (int j = A.row[i]; j < A.dia[i]; ++j) s = s - A.val[j] * x[A.col[j]];
  x[i] = s * d[i];
}
// x_i = (x_i / (d_i + TINY) - sum_{j>i} a_ij x_j) d_i
// (the scale pass Prec_Solve.f90:55-57 fused into the backward sweep :60-67)
__device__ __forceinline__ void bwd_row(const RowsCRS &A, const double *d, double *x, int i) {
  double s = x[i] / (d[i] + tfs::kTiny);
  for (int j = A.dia[i] + 1; j < A.row[i + 1]; ++j) s = s - A.val[j] * x[A.col[j]];
  x[i] = s * d[i];
}

enum SweepOp { OP_FACTOR = 0, OP_FWD, OP_BWD };

template <int OP>
__device__ __forceinline__ void sweep_row(const RowsCRS &A, double *d, const double *b, double *x,
                                          int i) {
  if (OP == OP_FACTOR) factor_row(A, d, i);
  if (OP == OP_FWD) fwd_row(A, d, b, x, i);
  if (OP == OP_BWD) bwd_row(A, d, x, i);
}

// One big level: rows perm[p0..p1).
template <int OP>
__global__ void __launch_bounds__(kLevelThreads)
k_level(int p0, int p1, const int *__restrict__ perm, RowsCRS A, double *d,
        const double *__restrict__ b, double *x, const KState *st) {
  if (st && tfs::solve_done(st)) return;
  const int p = p0 + blockIdx.x * kLevelThreads + threadIdx.x;
  if (p < p1) sweep_row<OP>(A, d, b, x, perm[p]);
}

// A run of small levels [l0, l1) walked by one workgroup.  Writes of one
// level are visible to the whole workgroup after the barrier.
template <int OP>
__global__ void __launch_bounds__(kSmallThreads)
k_levels_small(int l0, int l1, const int *__restrict__ lptr, const int *__restrict__ perm,
               RowsCRS A, double *d, const double *__restrict__ b, double *x, const KState *st) {
  if (st && tfs::solve_done(st)) return;
  for (int l = l0; l < l1; ++l) {
    const int p1 = lptr[l + 1];
    for (int p = lptr[l] + threadIdx.x; p < p1; p += kSmallThreads)
      sweep_row<OP>(A, d, b, x, perm[p]);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Host side.

struct Lib {
  std::mutex mu;
  bool ready = false;
  int device = 0;
  int n_cu = 0;
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

inline bool tm_active(int kc) { return g_tm.on && g_tm.kclass == kc && g_tm.used + 2 <= g_tm.ev.size(); }
inline void tm_begin(int kc) {
  if (tm_active(kc)) (void)hipEventRecord(g_tm.ev[g_tm.used], g_lib.stream);
}
inline void tm_end(int kc) {
  if (tm_active(kc)) {
    (void)hipEventRecord(g_tm.ev[g_tm.used + 1], g_lib.stream);
    g_tm.used += 2;
  }
}
// Call with the stream idle.
void tm_harvest() {
  for (size_t k = 0; k + 1 < g_tm.used; k += 2) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, g_tm.ev[k], g_tm.ev[k + 1]) == hipSuccess) {
      g_tm.ms += t;
      g_tm.launches++;
    }
  }
  g_tm.used = 0;
}

#define TFS_HIP(call)                                                               \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      std::fprintf(stderr, "tfs: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(e_), \
                   __FILE__, __LINE__);                                             \
      return e_ == hipErrorOutOfMemory ? TFS_ERR_NOMEM : TFS_ERR_DEVICE;            \
    }                                                                               \
  } while (0)

#define TFS_TRY(expr)            \
  do {                           \
    int rc_ = (expr);            \
    if (rc_ < 0) return rc_;     \
  } while (0)

struct Segment {
  int l0, l1;   // levels [l0, l1)
  int rows;
  bool small;
};

struct Schedule {
  int nlev = 0;
  std::vector<int> ptr;      // host copy of level starts
  std::vector<Segment> segs;
  int *d_perm = nullptr;
  int *d_ptr = nullptr;
};

}  // namespace

struct tfs_matrix {
  int n = 0;
  long long nnz = 0;
  int *d_row = nullptr, *d_col = nullptr, *d_dia = nullptr;
  double *d_val = nullptr;
  int nblk = 0;
  int *d_rblk = nullptr;
  Schedule lo, up;
  double *d_d = nullptr;       // factor: used diagonal of D
  int factor_prec = -1;        // prec the factor was built for, -1 = stale
  double *d_vec = nullptr;     // Krylov work vectors
  int n_vec = 0;
  double *d_x = nullptr, *d_r = nullptr;  // staging for host-memory solves
  KState *d_st = nullptr;
  KState *h_st = nullptr;      // pinned
  double *d_partials = nullptr;
  long long bytes = 0;
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

int ensure_init() {
  std::lock_guard<std::mutex> lk(g_lib.mu);
  if (g_lib.ready) return 0;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return TFS_ERR_NO_DEVICE;
  if (g_lib.device >= ndev) return TFS_ERR_INVALID;
  TFS_HIP(hipSetDevice(g_lib.device));
  TFS_HIP(hipDeviceGetAttribute(&g_lib.n_cu, hipDeviceAttributeMultiprocessorCount, g_lib.device));
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

// Level of each row in the lower (forward) or upper (backward) DAG; rows are
// then ordered by level, ascending row number within a level.
void build_schedule(int n, const std::vector<int> &row, const std::vector<int> &col,
                    const std::vector<int> &dia, bool lower, Schedule &S,
                    std::vector<int> &perm) {
  std::vector<int> lev(n, 0);
  int nlev = 0;
  if (lower) {
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
  S.ptr.assign(nlev + 1, 0);
  for (int i = 0; i < n; ++i) S.ptr[lev[i] + 1]++;
  for (int l = 0; l < nlev; ++l) S.ptr[l + 1] += S.ptr[l];
  perm.assign(n, 0);
  std::vector<int> fill(S.ptr.begin(), S.ptr.end() - 1);
  for (int i = 0; i < n; ++i) perm[fill[lev[i]]++] = i;
  S.segs.clear();
  for (int l = 0; l < nlev; ++l) {
    const int rows = S.ptr[l + 1] - S.ptr[l];
    const bool small = rows <= kSmallLevel;
    if (small && !S.segs.empty() && S.segs.back().small) {
      S.segs.back().l1 = l + 1;
      S.segs.back().rows += rows;
    } else {
      S.segs.push_back({l, l + 1, rows, small});
    }
  }
}

template <int OP>
void launch_sweep(const tfs_matrix *A, const Schedule &S, double *d, const double *b, double *x,
                  const KState *st) {
  const RowsCRS M{A->d_row, A->d_col, A->d_dia, A->d_val};
  const int kc = OP == OP_FACTOR ? TFS_KC_FACTOR : TFS_KC_SWEEP;
  for (const Segment &g : S.segs) {
    tm_begin(kc);
    if (g.small) {
      hipLaunchKernelGGL(k_levels_small<OP>, dim3(1), dim3(kSmallThreads), 0, g_lib.stream, g.l0,
                         g.l1, S.d_ptr, S.d_perm, M, d, b, x, st);
    } else {
      const int p0 = S.ptr[g.l0], p1 = S.ptr[g.l1];
      hipLaunchKernelGGL(k_level<OP>, dim3((p1 - p0 + kLevelThreads - 1) / kLevelThreads),
                         dim3(kLevelThreads), 0, g_lib.stream, p0, p1, S.d_perm, M, d, b, x, st);
    }
    tm_end(kc);
  }
}

int prec_form(tfs_matrix *A, int prec) {
  if (prec < 0 || prec > 2) prec = 0;  // the reference treats any other value as none
  if (A->factor_prec == prec) return 0;
  if (A->n > 0) {
    if (prec == 2) {
      launch_sweep<OP_FACTOR>(A, A->lo, A->d_d, nullptr, nullptr, nullptr);
    } else {
      hipLaunchKernelGGL(k_factor_diag, dim3(grid_for(A->n, kRedThreads * 4)), dim3(kRedThreads),
                         0, g_lib.stream, A->n, A->d_val, A->d_dia, A->d_d, prec);
    }
    TFS_HIP(hipGetLastError());
  }
  A->factor_prec = prec;
  return 0;
}

// x := M^-1 b.  For prec 0 the caller may alias instead of copying.
void prec_apply(const tfs_matrix *A, int prec, const double *b, double *x, const KState *st) {
  if (prec == 2) {
    launch_sweep<OP_FWD>(A, A->lo, A->d_d, b, x, st);
    launch_sweep<OP_BWD>(A, A->up, A->d_d, nullptr, x, st);
  } else if (prec == 1) {
    hipLaunchKernelGGL(k_jacobi, dim3(grid_for(A->n, kRedThreads * 4)), dim3(kRedThreads), 0,
                       g_lib.stream, A->n, b, A->d_d, x, st);
  } else {
    hipLaunchKernelGGL(k_copy, dim3(grid_for(A->n, kRedThreads * 4)), dim3(kRedThreads), 0,
                       g_lib.stream, A->n, b, x, st);
  }
}

template <int MODE>
void launch_spmv(const tfs_matrix *A, const double *x, double *y, const double *w, double *xs,
                 double *rs, KState *st) {
  tm_begin(TFS_KC_SPMV);
  hipLaunchKernelGGL(k_spmv<MODE>, dim3(grid_for(A->nblk, kSpmvWaves)), dim3(kSpmvThreads), 0,
                     g_lib.stream, A->nblk, A->d_rblk, A->d_row, A->d_col, A->d_val, x, y, w, xs,
                     rs, st, A->d_partials);
  tm_end(TFS_KC_SPMV);
}

int ensure_vectors(tfs_matrix *A, int count) {
  if (A->n_vec >= count) return 0;
  if (A->d_vec) {
    TFS_HIP(hipFree(A->d_vec));
    A->allocs.erase(std::find(A->allocs.begin(), A->allocs.end(), (void *)A->d_vec));
    A->bytes -= (long long)A->n_vec * A->n * (long long)sizeof(double);
    A->d_vec = nullptr;
  }
  TFS_TRY(dev_alloc(A, (void **)&A->d_vec, sizeof(double) * (size_t)A->n * count));
  A->n_vec = count;
  return 0;
}

inline int vec_grid(const tfs_matrix *A) { return grid_for(A->n, kRedThreads * 4); }

// One CG iteration (Cg.f90:97-178); Bicg takes the same path (its shadow
// sequence equals the primary one bit for bit, Bicg.f90:91-175).
void cg_iteration(tfs_matrix *A, int prec, double *x, double *r, double *z, double *p, double *q) {
  KState *st = A->d_st;
  const int n = A->n, g = vec_grid(A);
  double *zz = prec == 0 ? r : z;  // prec 0: z is a copy of r (Prec_Solve.f90:72-75)
  if (prec == 2) prec_apply(A, 2, r, z, st);
  hipLaunchKernelGGL(k_rho, dim3(g), dim3(kRedThreads), 0, g_lib.stream, n, r, zz, A->d_d,
                     prec == 1 ? 1 : 0, st, A->d_partials);
  hipLaunchKernelGGL(k_pupdate, dim3(g), dim3(kRedThreads), 0, g_lib.stream, n, zz, p, st);
  launch_spmv<SPMV_DOT>(A, p, q, p, nullptr, nullptr, st);
  hipLaunchKernelGGL(k_axpy_norm, dim3(g), dim3(kRedThreads), 0, g_lib.stream, n, x, r, p, q, st,
                     A->d_partials);
}

// One CGS iteration (Cgs.f90:104-226).
void cgs_iteration(tfs_matrix *A, int prec, double *x, double *r1, double *v) {
  KState *st = A->d_st;
  const int n = A->n, g = vec_grid(A);
  double *p1 = v, *p2 = v + n, *q1 = v + 2 * (size_t)n, *r2 = v + 3 * (size_t)n,
         *u1 = v + 4 * (size_t)n, *u2 = v + 5 * (size_t)n, *v2 = v + 6 * (size_t)n,
         *w = v + 7 * (size_t)n;
  // rho = (r1, r2)
  hipLaunchKernelGGL(k_rho, dim3(g), dim3(kRedThreads), 0, g_lib.stream, n, r1, r2, A->d_d, 0, st,
                     A->d_partials);
  hipLaunchKernelGGL(k_cgs_u, dim3(g), dim3(kRedThreads), 0, g_lib.stream, n, r1, q1, u1, u2, st);
  double *pp2 = prec == 0 ? u2 : p2;
  if (prec != 0) prec_apply(A, prec, u2, p2, st);
  launch_spmv<SPMV_DOT>(A, pp2, v2, r2, nullptr, nullptr, st);
  hipLaunchKernelGGL(k_cgs_q, dim3(g), dim3(kRedThreads), 0, g_lib.stream, n, u1, v2, q1, w, st);
  double *pp1 = prec == 0 ? w : p1;
  if (prec != 0) prec_apply(A, prec, w, p1, st);
  // q2 = A p1 into the v2 slot (v2 is dead after q1 and w were formed)
  launch_spmv<SPMV_CGS_TAIL>(A, pp1, v2, nullptr, x, r1, st);
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
  const int nvec = method == TFS_CGS ? 8 : 3;
  TFS_TRY(ensure_vectors(A, nvec));
  TFS_TRY(prec_form(A, prec));  // Cg.f90:42 (reused while the values are unchanged)

  KState *h = A->h_st;
  std::memset(h, 0, sizeof(KState));
  h->tol = tol;
  h->cap = cap;
  h->ini_res = *ini_res;
  TFS_HIP(hipMemcpyAsync(A->d_st, h, sizeof(KState), hipMemcpyHostToDevice, g_lib.stream));
  const int g = vec_grid(A);
  KState *st = A->d_st;
  // ||b|| and its early exit; r = b - A x; ||r|| and its early exit
  hipLaunchKernelGGL(k_sumsq, dim3(g), dim3(kRedThreads), 0, g_lib.stream, n, r1, (int)EPI_BNRM,
                     st, A->d_partials);
  launch_spmv<SPMV_RESID>(A, x, r1, nullptr, nullptr, nullptr, st);
  if (method == TFS_CGS)
    hipLaunchKernelGGL(k_copy, dim3(g), dim3(kRedThreads), 0, g_lib.stream, n, r1,
                       A->d_vec + 3 * (size_t)n, st);  // r2 = r1 (Cgs.f90:75-77)
  hipLaunchKernelGGL(k_sumsq, dim3(g), dim3(kRedThreads), 0, g_lib.stream, n, r1,
                     (int)EPI_INIRES, st, A->d_partials);
  TFS_HIP(hipGetLastError());
  TFS_TRY(read_state(A));

  if (!h->done) {
    if (cap < 1) {  // zero-trip DO loop: iter = 1 (Cg.f90:97,181)
      h->done = 1;
      h->niter_out = 1;
      h->fin_res = h->err;
      h->status = tfs::kStNotConverged;
    }
    double *v = A->d_vec;
    int queued = 0;
    int chunk = 1;
    double err_prev = h->err;
    int it_prev = 0;
    while (!h->done) {
      chunk = std::min(chunk, cap - queued);
      if (chunk < 1) chunk = 1;
      for (int k = 0; k < chunk; ++k) {
        if (method == TFS_CGS)
          cgs_iteration(A, prec, x, r1, v);
        else
          cg_iteration(A, prec, x, r1, v, v + n, v + 2 * (size_t)n);
      }
      queued += chunk;
      TFS_HIP(hipGetLastError());
      TFS_TRY(read_state(A));
      if (h->done) break;
      if (queued > cap + 64) return TFS_ERR_DEVICE;  // the norm kernel stops every solve at cap
      // Next chunk: iterations still needed at the observed contraction rate,
      // so that few iterations are queued past convergence.
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
  }
  *niter = h->niter_out;
  *fin_res = h->fin_res;
  *ini_res = h->ini_res;  // still the caller's value after the ||b|| < tol exit
  if (h->status == tfs::kStBreakdown) return TFS_ERR_BREAKDOWN;
  return h->status == tfs::kStNotConverged ? TFS_NOT_CONVERGED : TFS_OK;
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

  tfs_matrix *A = new tfs_matrix;
  A->n = n;
  A->nnz = nnz;
  auto fail = [&](int rc) {
    tfs_matrix_destroy(A);
    return rc;
  };
  int rc;
#define TFS_STEP(expr)          \
  if ((rc = (expr)) < 0) return fail(rc)

  // SpMV row blocks: <= 64 rows and <= kSpmvCap entries each.
  std::vector<int> rb;
  rb.push_back(0);
  for (int r = 0; r < n;) {
    int e = r + 1;
    while (e < n && e - r < tfs::kWave && hr[e + 1] - hr[r] <= kSpmvCap) ++e;
    rb.push_back(e);
    r = e;
  }
  A->nblk = (int)rb.size() - 1;

  std::vector<int> lperm, uperm;
  build_schedule(n, hr, hc, hd, true, A->lo, lperm);
  build_schedule(n, hr, hc, hd, false, A->up, uperm);

  TFS_STEP(dev_alloc(A, (void **)&A->d_row, sizeof(int) * (n + 1)));
  TFS_STEP(dev_alloc(A, (void **)&A->d_col, sizeof(int) * (size_t)nnz));
  TFS_STEP(dev_alloc(A, (void **)&A->d_dia, sizeof(int) * (size_t)n));
  TFS_STEP(dev_alloc(A, (void **)&A->d_val, sizeof(double) * (size_t)nnz));
  TFS_STEP(dev_alloc(A, (void **)&A->d_rblk, sizeof(int) * rb.size()));
  TFS_STEP(dev_alloc(A, (void **)&A->lo.d_perm, sizeof(int) * (size_t)n));
  TFS_STEP(dev_alloc(A, (void **)&A->lo.d_ptr, sizeof(int) * A->lo.ptr.size()));
  TFS_STEP(dev_alloc(A, (void **)&A->up.d_perm, sizeof(int) * (size_t)n));
  TFS_STEP(dev_alloc(A, (void **)&A->up.d_ptr, sizeof(int) * A->up.ptr.size()));
  TFS_STEP(dev_alloc(A, (void **)&A->d_d, sizeof(double) * (size_t)n));
  TFS_STEP(dev_alloc(A, (void **)&A->d_x, sizeof(double) * (size_t)n));
  TFS_STEP(dev_alloc(A, (void **)&A->d_r, sizeof(double) * (size_t)n));
  TFS_STEP(dev_alloc(A, (void **)&A->d_st, sizeof(KState)));
  TFS_STEP(dev_alloc(A, (void **)&A->d_partials, sizeof(double) * kMaxGrid));
  if (hipHostMalloc((void **)&A->h_st, sizeof(KState), hipHostMallocDefault) != hipSuccess)
    return fail(TFS_ERR_NOMEM);
#undef TFS_STEP
  auto up = [&](void *dst, const void *src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  bool ok = up(A->d_row, hr.data(), sizeof(int) * (n + 1)) &&
            up(A->d_col, hc.data(), sizeof(int) * (size_t)nnz) &&
            up(A->d_dia, hd.data(), sizeof(int) * (size_t)n) &&
            up(A->d_rblk, rb.data(), sizeof(int) * rb.size()) &&
            up(A->lo.d_perm, lperm.data(), sizeof(int) * (size_t)n) &&
            up(A->lo.d_ptr, A->lo.ptr.data(), sizeof(int) * A->lo.ptr.size()) &&
            up(A->up.d_perm, uperm.data(), sizeof(int) * (size_t)n) &&
            up(A->up.d_ptr, A->up.ptr.data(), sizeof(int) * A->up.ptr.size()) &&
            hipMemset(A->d_st, 0, sizeof(KState)) == hipSuccess &&
            hipMemset(A->d_val, 0, sizeof(double) * (size_t)nnz) == hipSuccess;
  if (!ok) return fail(TFS_ERR_DEVICE);
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
  TFS_HIP(hipMemcpyAsync(A->d_val, val, sizeof(double) * (size_t)A->nnz, hipMemcpyHostToDevice,
                         g_lib.stream));
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  A->factor_prec = -1;
  return TFS_OK;
}

extern "C" int tfs_matrix_set_values_device(tfs_matrix *A, const double *d_val) {
  if (!A || (A->nnz > 0 && !d_val)) return TFS_ERR_INVALID;
  TFS_HIP(hipMemcpyAsync(A->d_val, d_val, sizeof(double) * (size_t)A->nnz,
                         hipMemcpyDeviceToDevice, g_lib.stream));
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  A->factor_prec = -1;
  return TFS_OK;
}

extern "C" int tfs_matrix_values_changed(tfs_matrix *A) {
  if (!A) return TFS_ERR_INVALID;
  A->factor_prec = -1;
  return TFS_OK;
}

extern "C" double *tfs_matrix_values_device(tfs_matrix *A) { return A ? A->d_val : nullptr; }

extern "C" int tfs_matrix_get_info(const tfs_matrix *A, tfs_matrix_info *info) {
  if (!A || !info) return TFS_ERR_INVALID;
  info->n = A->n;
  info->nnz = A->nnz;
  info->levels_lower = A->lo.nlev;
  info->levels_upper = A->up.nlev;
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
    TFS_HIP(hipMemcpyAsync(A->d_x, x1, bytes, hipMemcpyHostToDevice, g_lib.stream));
    TFS_HIP(hipMemcpyAsync(A->d_r, r1, bytes, hipMemcpyHostToDevice, g_lib.stream));
  }
  const int rc = solve_device(A, method, prec, A->d_x, A->d_r, niter, tol, ini_res, fin_res);
  if (rc < 0 && rc != TFS_ERR_BREAKDOWN) return rc;
  if (A->n > 0) {
    TFS_HIP(hipMemcpyAsync(x1, A->d_x, bytes, hipMemcpyDeviceToHost, g_lib.stream));
    TFS_HIP(hipMemcpyAsync(r1, A->d_r, bytes, hipMemcpyDeviceToHost, g_lib.stream));
    TFS_HIP(hipStreamSynchronize(g_lib.stream));
  }
  return rc;
}

extern "C" int tfs_spmv_device(tfs_matrix *A, const double *d_x, double *d_y) {
  if (!A || (A->n > 0 && (!d_x || !d_y))) return TFS_ERR_INVALID;
  if (A->n == 0) return TFS_OK;
  launch_spmv<SPMV_PLAIN>(A, d_x, d_y, nullptr, nullptr, nullptr, nullptr);
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  return TFS_OK;
}

extern "C" int tfs_residual_device(tfs_matrix *A, const double *d_x, double *d_r) {
  if (!A || (A->n > 0 && (!d_x || !d_r))) return TFS_ERR_INVALID;
  if (A->n == 0) return TFS_OK;
  launch_spmv<SPMV_RESID>(A, d_x, d_r, nullptr, nullptr, nullptr, nullptr);
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  return TFS_OK;
}

extern "C" int tfs_prec_form(tfs_matrix *A, int prec) {
  if (!A) return TFS_ERR_INVALID;
  TFS_TRY(prec_form(A, prec));
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  return TFS_OK;
}

extern "C" int tfs_prec_solve_device(tfs_matrix *A, int prec, const double *d_b, double *d_x) {
  if (!A || (A->n > 0 && (!d_b || !d_x))) return TFS_ERR_INVALID;
  if (prec < 0 || prec > 2) prec = 0;
  if (A->factor_prec != prec) return TFS_ERR_INVALID;  // tfs_prec_form first
  if (A->n == 0) return TFS_OK;
  prec_apply(A, prec, d_b, d_x, nullptr);
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  return TFS_OK;
}

extern "C" int tfs_prec_diag_device(tfs_matrix *A, double *d_out) {
  if (!A || (A->n > 0 && !d_out)) return TFS_ERR_INVALID;
  if (A->factor_prec < 0) return TFS_ERR_INVALID;
  TFS_HIP(hipMemcpyAsync(d_out, A->d_d, sizeof(double) * (size_t)A->n, hipMemcpyDeviceToDevice,
                         g_lib.stream));
  TFS_HIP(hipStreamSynchronize(g_lib.stream));
  return TFS_OK;
}

extern "C" int tfs_spmv_time(tfs_matrix *A, const double *d_x, double *d_y, int reps, double *ms) {
  if (!A || !ms || reps < 1 || (A->n > 0 && (!d_x || !d_y))) return TFS_ERR_INVALID;
  if (A->n == 0) {
    *ms = 0.0;
    return TFS_OK;
  }
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
