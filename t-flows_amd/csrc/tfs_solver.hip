// tfs_solver.hip -- host side of the MI355X (gfx950) drop-in for T-FlowS's
// Krylov solvers, and its C ABI (include/tfs_solver.h).
//
// Data layout (tfs_layout.hpp): the matrix lives in three row orders.  S
// (reverse Cuthill-McKee) carries the SpMV and every Krylov vector; L and U
// carry the lower / upper triangles level by level for the factor and the two
// sweeps.  Rows keep the caller's entry order, so every row sum is the
// reference's; the caller's x and b are permuted in and out once per solve.
// Values exist in FP64 (the drop-in) and, on demand, FP32 (the inner solves
// of iterative refinement).  Kernels: tfs_kernels.hpp.
//
// A solve queues its iterations without host round trips: the scalars live on
// the device (KState) and every kernel returns at once after convergence; the
// host only reads the state between chunks of iterations, sizing each chunk
// from the observed contraction rate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "tfs_kernels.hpp"
#include "tfs_layout.hpp"
#include "tfs_solver.h"
#include "tfs_internal.hpp"

using namespace tfs;

// Scratch of tfs_residual_device_noalloc: the two S-order vectors of the
// residual (x, r), up to kScratchRows rows each -- 64 MiB held on every GPU
// the library runs on until the process ends.  Host code must not pass
// `scratch` itself to a kernel (on the host it names a stand-in, not device
// memory): its device address is looked up once per device at set-up
// (ensure_init) and kept in Lib::scratch.
constexpr int kScratchRows = 1 << 22;
static __device__ __align__(256) double scratch[2 * (size_t)kScratchRows];

namespace {

constexpr int kMaxDevices = 64;

struct Lib {
  std::mutex mu;
  bool ready = false;
  int device = 0;
  double *scratch[kMaxDevices] = {};  // device address of `scratch`, per device
};
Lib g_lib;

// Every matrix (and batch) owns a HIP stream; an API call runs all its work
// on the stream of the matrix it was given, so solves of different matrices
// issued from different host threads overlap on the device.
thread_local hipStream_t t_stream = nullptr;
inline hipStream_t cur_stream() { return t_stream; }
struct StreamScope {
  hipStream_t prev;
  explicit StreamScope(hipStream_t s) : prev(t_stream) { t_stream = s; }
  ~StreamScope() { t_stream = prev; }
};

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
  if (tm_active(kc)) (void)hipEventRecord(g_tm.ev[g_tm.used], cur_stream());
}
inline void tm_end(int kc) {
  if (tm_active(kc)) {
    (void)hipEventRecord(g_tm.ev[g_tm.used + 1], cur_stream());
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
  std::vector<Segment> segs;
  int *lev = nullptr, *ptr = nullptr, *col = nullptr;
  int *of_orig = nullptr;
  long long ntri = 0;
};

// Values and factor in one precision, in the S, L and U layouts.
template <typename T>
struct Values {
  T *s_val = nullptr, *s_diag = nullptr, *l_val = nullptr, *u_val = nullptr, *l_diag = nullptr;
  T *dL = nullptr, *dU = nullptr;
  bool allocated = false;
  bool ready = false;     // copies match the caller's values
  int factor_prec = -1;   // prec the factor was built for, -1 = stale
};

// Krylov work vectors of one precision (S order unless noted).
template <typename T>
struct Vecs {
  T *x, *r, *z, *p, *q, *yL /* L order */, *zU /* U order */;
  // CGS only (FP64)
  T *p1, *p2, *q1, *r2, *u1, *u2, *v2, *w;
};

}  // namespace

struct tfs_matrix {
  int n = 0;
  long long nnz = 0;
  // S layout
  int *s_row = nullptr, *s_col = nullptr, *s_src = nullptr, *s_diag_src = nullptr;
  int *s_orig = nullptr, *s_of_orig = nullptr, *rblk = nullptr;
  int nblk = 0;
  int spmv_rows = kSpmvRows;    // rows per SpMV row block (one wavefront each)
  int spmv_waves = kSpmvWaves;  // wavefronts per SpMV workgroup
  DevSweep L, U;
  int *l_to_s = nullptr, *u_to_l = nullptr, *s_to_u = nullptr;
  // permutations that are the identity or the reversal, formed in the
  // kernels instead of loaded (the U order is the L order backwards; in a
  // level-sorted copy the L order is the S order)
  bool l_ident = false, u_rev = false, s_to_u_rev = false;
  int *s_lpos = nullptr;  // L position of the row at S position k
  double *val = nullptr;       // caller-order FP64 values
  Values<double> v64;
  Values<float> v32;
  double *vec64 = nullptr;     // FP64 work vectors
  int n_vec64 = 0;
  float *vec32 = nullptr;      // FP32 work vectors (iterative refinement)
  int n_vec32 = 0;
  double *stage_x = nullptr, *stage_r = nullptr;  // host-memory solves
  KState *d_st = nullptr;      // [0] Krylov solve, [1] refinement outer loop
  KState *h_st = nullptr;      // pinned, 2 entries
  double *partials = nullptr;
  int *sweep_err = nullptr;    // set by a persistent sweep whose wait ran out
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool persistent = true;      // one launch per sweep (else one per level)
  bool dynamic = false;        // persistent sweeps take chunks from a counter
  bool coop = false;           // static schedule under a cooperative launch
  bool sweep_fixed = false;    // schedule set by the caller or the environment
  bool sweep_tuned = false;    // schedule picked by timing (first factor)
  unsigned int *sweep_ctr = nullptr;  // the dynamic schedule's counters
  int grid_persistent = 0;
  bool grid_fixed = false;     // grid limited by the caller or the environment
  int sweep_capacity = 1;      // co-resident workgroups of the heaviest sweep kernel
  long long bytes = 0;
  int bandwidth = 0;
  std::vector<void *> allocs;
  // FP32 inner solves of the refinement on a colour-ordered copy P A P^T,
  // P a greedy colouring of the S-order graph (built on first use)
  int inner_order = TFS_INNER_AUTO;
  int inner_built = -1;           // the order the inner copy was built in
  int inner_brick = 0;            // its brick edge (bricks only)
  int inner_block = kInnerBlock;  // rows coloured as one unit in the inner order
  int refine = TFS_REFINE_RELIABLE;  // outer loop of the refinement
  tfs_matrix *inner = nullptr;
  int *inner_emap = nullptr;     // inner's caller-order entry k <- entry of val
  int *inner_from_s = nullptr;   // S position of this matrix for inner's S position i
  bool inner_ready = false;      // inner's values match val
  int inner_colours = 0;
  // The FP32 operator of the reliable-update refinement (on the inner copy):
  // sliced ELL of the off-diagonals in S order plus FP32 row sums.
  int ell_nslice = 0;
  long long ell_len = 0;
  long long *ell_off = nullptr;
  int *ell_col = nullptr, *ell_src = nullptr;
  float *ell_val = nullptr, *ell_rsum = nullptr;
  bool ell_ready = false;
  // FP64 SpMV over a sliced ELL of all entries in S order (k_ell64); absent
  // (e64_len 0) when padding would cost more than a quarter of nnz
  int e64_nslice = 0;
  long long e64_len = 0;
  long long *e64_off = nullptr;
  int *e64_col = nullptr, *e64_src = nullptr;
  double *e64_val = nullptr;
  // Point-coloured incomplete-Cholesky sweeps of the inner copy: per colour,
  // sliced ELL of the strictly-lower (L) and strictly-upper (U) entries.
  int csw_colours = 0;
  std::vector<int> csw_cstart, csw_sbeg;  // host: first row / first slice of each colour
  int *csw_row0 = nullptr;                // first row of each slice
  long long csw_len[2] = {0, 0};
  long long *csw_off[2] = {nullptr, nullptr};
  int *csw_col[2] = {nullptr, nullptr}, *csw_src[2] = {nullptr, nullptr};
  float *csw_val[2] = {nullptr, nullptr};
  bool csw_ready = false;
};

namespace {

int matrix_create(int n, long long nnz, const int *row, const int *col, const int *dia, int base,
                  tfs_matrix **out, Layout *keep, bool identity_s = false);

int dev_alloc(tfs_matrix *A, void **p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  TFS_HIP(hipMalloc(p, bytes));
  A->allocs.push_back(*p);
  A->bytes += (long long)bytes;
  return 0;
}

int dev_free(tfs_matrix *A, void *p, size_t bytes) {
  if (!p) return 0;
  TFS_HIP(hipFree(p));
  A->allocs.erase(std::find(A->allocs.begin(), A->allocs.end(), p));
  A->bytes -= (long long)bytes;
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
  if (g_lib.device >= ndev || g_lib.device >= kMaxDevices) return TFS_ERR_INVALID;
  TFS_HIP(hipSetDevice(g_lib.device));
  // this device's copy of the allocation-free entry point's scratch array
  void *sp = nullptr;
  TFS_HIP(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
  g_lib.scratch[g_lib.device] = static_cast<double *>(sp);
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
inline int ell64_grid(const tfs_matrix *A) { return grid_for(A->e64_nslice, kEllWaves); }
// Slice walk of the ELL SpMVs (SliceWalk): measured at 256^3 / 128^3, the
// XCD-contiguous walk is faster for the FP32 operator in colour order
// (1022 against 1039 ms of SpMV per solve; 60 against 65) and slower for the
// FP64 one in RCM order (719 against 662; 42 against 41).
constexpr int kEll64Xcd = 0, kEllDiffXcd = 1;
// TFS_INNER_AUTO: bricks from this size (measured on MI355X, Poisson box of
// aspect 100, point colouring against 4^3 bricks per solve: 128^3 217 / 215
// ms, 192^3 985 / 818 ms, 256^3 2997 / 2484 ms)
constexpr int kBrickAutoRows = 1 << 22;
// The brick edge grows with the system: larger bricks are a stronger factor
// (fewer FP32 iterations) whose sweeps have more, narrower levels, and a level
// costs a dependency round trip however narrow it is; the edge is the largest
// whose levels keep about 2e5 rows or more (on a lexicographic box a two-colour
// brick order of edge E has 4E - 2 levels).  Measured per solve, bricks of
// 4^3 / 8^3 / 16^3: 192^3 818 / 733 / 803 ms, 256^3 2484 / 2167 / 2097 ms
// (and 6^3 2365, 12^3 2211, 32^3 2428 ms at 256^3).
constexpr int kBrick8Rows = 3 << 21, kBrick16Rows = 3 << 22;
inline int brick_edge(int n) { return n >= kBrick16Rows ? 16 : n >= kBrick8Rows ? 8 : 4; }
// Partial sums a fused-dot SpMV of precision T leaves (one per workgroup).
template <typename T>
inline int spmv_parts(const tfs_matrix *A) {
  if (sizeof(T) == 8 && A->e64_len > 0) return ell64_grid(A);
  return grid_for(A->nblk, A->spmv_waves);
}

template <typename TO, typename TI>
void gather(int n, const TI *in, const int *idx, TO *out, const KState *st = nullptr) {
  hipLaunchKernelGGL((k_gather<TO, TI>), dim3(vgrid(n)), dim3(kRedThreads), 0, cur_stream(), n,
                     in, idx, out, st);
}
template <typename T>
void scatter(int n, const T *in, const int *idx, T *out) {
  hipLaunchKernelGGL(k_scatter<T>, dim3(vgrid(n)), dim3(kRedThreads), 0, cur_stream(), n, in, idx,
                     out);
}
template <typename T>
void arm(int n, T *x) {
  hipLaunchKernelGGL(k_arm<T>, dim3(vgrid(n)), dim3(kRedThreads), 0, cur_stream(), n, x);
}
void finish(int nparts, const double *partials, int epi, KState *st) {
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(kRedThreads), 0, cur_stream(), nparts, partials, epi,
                     st);
}
template <typename T>
Values<T> &vals(tfs_matrix *A);
template <>
Values<double> &vals<double>(tfs_matrix *A) { return A->v64; }
template <>
Values<float> &vals<float>(tfs_matrix *A) { return A->v32; }

template <typename T, int OP>
void launch_sweep(tfs_matrix *A, const DevSweep &S, const Tri<T> &Tr, T *d, const T *b, T *x,
                  const KState *st, T *armv = nullptr) {
  const int kc = OP == OP_FACTOR ? TFS_KC_FACTOR : TFS_KC_SWEEP;
  if (A->persistent && A->dynamic) {
    tm_begin(kc);
    hipLaunchKernelGGL((k_sweep_syncfree<T, OP, true>), dim3(A->grid_persistent), dim3(kChunk), 0,
                       cur_stream(), A->n, Tr, d, const_cast<T *>(b), x, armv, st, A->sweep_err,
                       A->sweep_ctr);
    tm_end(kc);
    return;
  }
  if (A->persistent && !A->coop) {
    // The static schedule needs every workgroup of the grid resident at
    // once.  The grid is capped at the occupancy of the heaviest sweep
    // kernel, so an ordinary launch has all its workgroups dispatched as
    // soon as the workgroups of other, terminating kernels have left; it
    // saves the 10-20 us of queue synchronisation a cooperative launch
    // costs on each side.  Only another never-ending kernel holding CUs
    // could keep a workgroup out: the bounded polls then end the sweep with
    // an error (solves sharing the GPU use the dynamic schedule).
    tm_begin(kc);
    hipLaunchKernelGGL((k_sweep_syncfree<T, OP, false>), dim3(A->grid_persistent), dim3(kChunk), 0,
                       cur_stream(), A->n, Tr, d, const_cast<T *>(b), x, armv, st, A->sweep_err,
                       A->sweep_ctr);
    tm_end(kc);
    return;
  }
  if (A->persistent) {
    tm_begin(kc);
    int nn = A->n;
    Tri<T> TT = Tr;
    T *dd = d, *bb = const_cast<T *>(b), *xx = x, *aa = armv;
    const KState *ss = st;
    int *ee = A->sweep_err;
    unsigned int *cc = A->sweep_ctr;
    void *args[] = {&nn, &TT, &dd, &bb, &xx, &aa, &ss, &ee, &cc};
    const hipError_t e = hipLaunchCooperativeKernel(
        reinterpret_cast<const void *>(k_sweep_syncfree<T, OP, false>), dim3(A->grid_persistent),
        dim3(kChunk), args, 0, cur_stream());
    tm_end(kc);
    if (e == hipSuccess) return;
    // No cooperative launch here: fall back to one launch per level for this
    // matrix from now on (those kernels need no sentinels).
    (void)hipGetLastError();
    std::fprintf(stderr, "tfs: cooperative launch failed (%s); using per-level sweeps\n",
                 hipGetErrorString(e));
    A->persistent = false;
  }
  // one launch per level (or run of small levels): the timing brackets the
  // whole sweep, so that a measured "launch" is one sweep in every schedule
  tm_begin(kc);
  for (const Segment &g : S.segs) {
    if (g.small) {
      hipLaunchKernelGGL((k_levels_small<T, OP>), dim3(1), dim3(kSmallThreads), 0, cur_stream(),
                         g.l0, g.l1, S.lev, Tr, d, b, x, st);
    } else {
      const int p0 = S.lev_ptr[g.l0], p1 = S.lev_ptr[g.l1];
      hipLaunchKernelGGL((k_level<T, OP>), dim3((p1 - p0 + kLevelThreads - 1) / kLevelThreads),
                         dim3(kLevelThreads), 0, cur_stream(), p0, p1, Tr, d, b, x, st);
    }
  }
  tm_end(kc);
}

template <typename T>
int alloc_values(tfs_matrix *A) {
  Values<T> &V = vals<T>(A);
  if (V.allocated) return 0;
  const size_t n = (size_t)std::max(A->n, 1);
  TFS_TRY(dev_alloc(A, (void **)&V.s_val, sizeof(T) * (size_t)std::max(A->nnz, 1LL)));
  TFS_TRY(dev_alloc(A, (void **)&V.l_val, sizeof(T) * (size_t)std::max(A->L.ntri, 1LL)));
  TFS_TRY(dev_alloc(A, (void **)&V.u_val, sizeof(T) * (size_t)std::max(A->U.ntri, 1LL)));
  TFS_TRY(dev_alloc(A, (void **)&V.s_diag, sizeof(T) * n));
  TFS_TRY(dev_alloc(A, (void **)&V.l_diag, sizeof(T) * n));
  TFS_TRY(dev_alloc(A, (void **)&V.dL, sizeof(T) * n));
  TFS_TRY(dev_alloc(A, (void **)&V.dU, sizeof(T) * n));
  V.allocated = true;
  return 0;
}

// Copies of the caller's values into the S, L and U layouts in precision T
// (part of the work the reference does each call: the matrix changed).
template <typename T>
int prepare_values(tfs_matrix *A) {
  TFS_TRY(alloc_values<T>(A));
  Values<T> &V = vals<T>(A);
  if (V.ready) return 0;
  const int n = A->n;
  if (n > 0)
    hipLaunchKernelGGL(k_values<T>, dim3(vgrid(n)), dim3(kRedThreads), 0, cur_stream(), n,
                       A->val, A->s_row, A->s_src, A->s_diag_src, A->s_lpos, A->s_to_u,
                       A->L.ptr, A->U.ptr, V.s_val, V.s_diag, V.l_val, V.u_val, V.l_diag);
  if (sizeof(T) == 8 && A->e64_len > 0)
    hipLaunchKernelGGL(k_ell_values<double>, dim3(grid_for(A->e64_len, kRedThreads, 8192)),
                       dim3(kRedThreads), 0, cur_stream(), A->e64_len, A->e64_src, A->val,
                       A->e64_val);
  V.ready = true;
  return 0;
}

void values_changed(tfs_matrix *A) {
  A->v64.ready = A->v32.ready = false;
  A->v64.factor_prec = A->v32.factor_prec = -1;
  A->inner_ready = false;
  A->ell_ready = false;
  A->csw_ready = false;
}

template <typename T>
int tune_sweeps(tfs_matrix *A);
int check_sweep(tfs_matrix *A);

template <typename T>
int prec_form(tfs_matrix *A, int prec) {
  if (prec < 0 || prec > 2) prec = 0;  // the reference treats any other value as none
  TFS_TRY(prepare_values<T>(A));
  Values<T> &V = vals<T>(A);
  if (V.factor_prec == prec) return 0;
  if (A->n > 0 && prec == 2) {
    const Tri<T> Tr{A->L.ptr, A->L.col, V.l_val, V.l_diag, nullptr};
    if (A->persistent) arm(A->n, V.dL);
    launch_sweep<T, OP_FACTOR>(A, A->L, Tr, V.dL, nullptr, nullptr, nullptr);
    gather(A->n, V.dL, A->u_to_l, V.dU);  // the factor in U order
    TFS_HIP(hipGetLastError());
  }
  V.factor_prec = prec;
  if (A->n > 0 && prec == 2 && !A->sweep_tuned && !A->sweep_fixed && A->persistent)
    TFS_TRY(tune_sweeps<T>(A));
  return 0;
}

// zU := M^-1 b for prec 2, b in S order, result in U order (yL scratch).
template <typename T>
void dic_apply(tfs_matrix *A, const T *bS, T *yL, T *zU, const KState *st) {
  Values<T> &V = vals<T>(A);
  // The FP32 sweeps form an identity or reversed right-hand-side index
  // instead of loading it; the FP64 ones (latency-bound) measured faster
  // loading it (256^3: 1430.5 against 1438.3 us per sweep; 128^3: 633.9
  // against 640.9).
  const bool form = sizeof(T) == sizeof(float);
  const Tri<T> TL{A->L.ptr, A->L.col, V.l_val, nullptr,
                  form && A->l_ident ? nullptr : A->l_to_s, 0};
  launch_sweep<T, OP_FWD>(A, A->L, TL, V.dL, bS, yL, st, zU);
  const Tri<T> TU{A->U.ptr, A->U.col, V.u_val, nullptr,
                  form && A->u_rev ? nullptr : A->u_to_l, form && A->u_rev ? A->n : 0};
  launch_sweep<T, OP_BWD>(A, A->U, TU, V.dU, yL, zU, st);
}

// xS := M^-1 bS, both in S order (CGS and the building block).
template <typename T>
void prec_apply_S(tfs_matrix *A, int prec, const T *bS, T *xS, T *yL, T *zU, const KState *st) {
  if (prec == 2) {
    dic_apply<T>(A, bS, yL, zU, st);
    gather(A->n, zU, A->s_to_u, xS, st);
  } else {
    hipLaunchKernelGGL(k_diag_apply<T>, dim3(vgrid(A->n)), dim3(kRedThreads), 0, cur_stream(),
                       A->n, bS, vals<T>(A).s_diag, xS, prec, st);
  }
}

template <typename T, int MODE>
void launch_spmv(tfs_matrix *A, const T *x, T *y, const T *w, T *xs, T *rs, const KState *st) {
  tm_begin(TFS_KC_SPMV);
  if constexpr (sizeof(T) == 8) {
    if (A->e64_len > 0) {
      hipLaunchKernelGGL(k_ell64<MODE>, dim3(ell64_grid(A)), dim3(kEllWaves * kWave), 0,
                         cur_stream(), A->n, A->e64_nslice, A->e64_off, A->e64_col, A->e64_val,
                         x, y, w, xs, rs, st, A->partials, kEll64Xcd);
      tm_end(TFS_KC_SPMV);
      return;
    }
  }
  const int g = spmv_parts<T>(A);
  const T *v = vals<T>(A).s_val;
  switch (A->spmv_waves) {
    case 2:
      hipLaunchKernelGGL((k_spmv<T, MODE, 2>), dim3(g), dim3(2 * kWave), 0, cur_stream(), A->nblk,
                         A->rblk, A->s_row, A->s_col, v, x, y, w, xs, rs, st, A->partials);
      break;
    case 8:
      hipLaunchKernelGGL((k_spmv<T, MODE, 8>), dim3(g), dim3(8 * kWave), 0, cur_stream(), A->nblk,
                         A->rblk, A->s_row, A->s_col, v, x, y, w, xs, rs, st, A->partials);
      break;
    default:
      hipLaunchKernelGGL((k_spmv<T, MODE, 4>), dim3(g), dim3(4 * kWave), 0, cur_stream(), A->nblk,
                         A->rblk, A->s_row, A->s_col, v, x, y, w, xs, rs, st, A->partials);
  }
  tm_end(TFS_KC_SPMV);
}

template <typename T>
Vecs<T> vecs_of(T *v, int n_) {
  const size_t n = (size_t)n_;
  Vecs<T> V;
  V.x = v; V.r = v + n; V.yL = v + 2 * n; V.zU = v + 3 * n;
  V.z = v + 4 * n; V.p = v + 5 * n; V.q = v + 6 * n;
  V.p1 = v + 4 * n; V.p2 = v + 5 * n; V.q1 = v + 6 * n; V.r2 = v + 7 * n;
  V.u1 = v + 8 * n; V.u2 = v + 9 * n; V.v2 = v + 10 * n; V.w = v + 11 * n;
  return V;
}

// (Re)allocates `count` work vectors of type T; slot 2 (the forward sweep's
// output) starts armed.
template <typename T>
int ensure_vectors(tfs_matrix *A, T **vec, int *have, int count) {
  if (*have >= count) return 0;
  TFS_TRY(dev_free(A, *vec, sizeof(T) * (size_t)A->n * (size_t)*have));
  *vec = nullptr;
  *have = 0;
  TFS_TRY(dev_alloc(A, (void **)vec, sizeof(T) * (size_t)A->n * count));
  *have = count;
  arm(A->n, *vec + 2 * (size_t)A->n);
  TFS_HIP(hipGetLastError());
  return 0;
}
int ensure_vec64(tfs_matrix *A, int count) { return ensure_vectors(A, &A->vec64, &A->n_vec64, count); }
int ensure_vec32(tfs_matrix *A, int count) { return ensure_vectors(A, &A->vec32, &A->n_vec32, count); }

template <typename T> T **vec_slots(tfs_matrix *A);
template <> double **vec_slots<double>(tfs_matrix *A) { return &A->vec64; }
template <> float **vec_slots<float>(tfs_matrix *A) { return &A->vec32; }
template <typename T> int *vec_count(tfs_matrix *A);
template <> int *vec_count<double>(tfs_matrix *A) { return &A->n_vec64; }
template <> int *vec_count<float>(tfs_matrix *A) { return &A->n_vec32; }

// Every sweep schedule gives the same bits; on the first factor of a matrix
// the persistent synchronisation-free sweeps (at the default grid and, unless
// the caller limited it, at a wider one) and the per-level launches are timed
// on it (two forward/backward pairs each) and the fastest is kept.  Deep DAGs
// of small matrices favour the levels (whose small levels run in one
// workgroup); deep wide ones the persistent kernel at the default grid, whose
// pollers then delay the stores they wait for the least; shallow DAGs with
// levels of a million rows (the 4^3 brick order of a 256^3 box: 14 levels)
// the persistent kernel at the wider grid (on MI355X 227 us a sweep with 1536
// workgroups, 332 with 768, 249 by per-level launches).
template <typename T>
int tune_sweeps(tfs_matrix *A) {
  A->sweep_tuned = true;
  TFS_TRY(ensure_vectors(A, vec_slots<T>(A), vec_count<T>(A), 7));
  const Vecs<T> V = vecs_of(*vec_slots<T>(A), A->n);
  hipLaunchKernelGGL(k_fill<T>, dim3(vgrid(A->n)), dim3(kRedThreads), 0, cur_stream(), A->n, V.z,
                     T(1));
  const int chunks = (A->n + kChunk - 1) / kChunk;
  const int grid[3] = {
      A->grid_persistent,
      A->grid_fixed ? A->grid_persistent
                    : std::max(1, std::min({chunks, A->sweep_capacity, kSweepGroupsWide})),
      A->grid_persistent};
  float t[3] = {0.f, 0.f, 0.f};
  for (int mode = 0; mode < 3; ++mode) {  // 0 persistent (default grid), 1 wider, 2 levels
    if (mode == 1 && grid[1] <= grid[0]) {
      t[1] = t[0];
      continue;
    }
    A->persistent = mode < 2;
    A->grid_persistent = grid[mode];
    dic_apply<T>(A, V.z, V.yL, V.zU, nullptr);  // warm
    TFS_HIP(hipEventRecord(A->ev0, cur_stream()));
    for (int k = 0; k < 2; ++k) dic_apply<T>(A, V.z, V.yL, V.zU, nullptr);
    TFS_HIP(hipEventRecord(A->ev1, cur_stream()));
    TFS_HIP(hipGetLastError());
    TFS_HIP(hipEventSynchronize(A->ev1));
    TFS_HIP(hipEventElapsedTime(&t[mode], A->ev0, A->ev1));
  }
  const int best = t[0] <= t[1] && t[0] <= t[2] ? 0 : t[1] <= t[2] ? 1 : 2;
  A->persistent = best < 2;
  A->grid_persistent = grid[best];
  if (A->persistent) {
    // the level launches left values where the persistent sweeps expect
    // sentinels: re-arm the forward sweep's output in both precisions
    if (A->vec64) arm(A->n, A->vec64 + 2 * (size_t)A->n);
    if (A->vec32) arm(A->n, A->vec32 + 2 * (size_t)A->n);
  }
  TFS_HIP(hipGetLastError());
  return check_sweep(A);
}

// One CG iteration (Cg.f90:97-178) in precision T; Bicg takes the same path
// (its shadow sequence equals the primary one bit for bit, Bicg.f90:91-175).
template <typename T>
void cg_iteration(tfs_matrix *A, int prec, const Vecs<T> &V, KState *st) {
  const int n = A->n, g = vgrid(n);
  if (prec == 2) dic_apply<T>(A, V.r, V.yL, V.zU, st);
  const int mode = prec == 2 ? (A->s_to_u_rev ? 4 : 2) : prec == 1 ? 1 : 0;
  T *z = prec == 0 ? V.r : V.z;  // prec 0: z is a copy of r (Prec_Solve.f90:72-75)
  hipLaunchKernelGGL(k_rho<T>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, V.r, V.z,
                     vals<T>(A).s_diag, V.zU, A->s_to_u, mode, st, A->partials);
  finish(g, A->partials, EPI_RHO, st);
  hipLaunchKernelGGL(k_pupdate<T>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, z, V.p, st);
  launch_spmv<T, SPMV_DOT>(A, V.p, V.q, V.p, nullptr, nullptr, st);
  finish(spmv_parts<T>(A), A->partials, EPI_ALPHA, st);
  hipLaunchKernelGGL(k_axpy_norm<T>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, V.x, V.r,
                     V.p, V.q, st, A->partials);
  finish(g, A->partials, EPI_NORM, st);
}

// One CGS iteration (Cgs.f90:104-226), FP64.
void cgs_iteration(tfs_matrix *A, int prec, const Vecs<double> &V, KState *st) {
  const int n = A->n, g = vgrid(n);
  hipLaunchKernelGGL(k_rho<double>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, V.r, V.r2,
                     nullptr, nullptr, nullptr, 3, st, A->partials);  // rho = (r1, r2)
  finish(g, A->partials, EPI_RHO, st);
  hipLaunchKernelGGL(k_cgs_u, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, V.r, V.q1, V.u1,
                     V.u2, st);
  double *pp2 = V.u2;  // prec 0: p2 is a copy of u2
  if (prec != 0) {
    prec_apply_S<double>(A, prec, V.u2, V.p2, V.yL, V.zU, st);
    pp2 = V.p2;
  }
  launch_spmv<double, SPMV_DOT>(A, pp2, V.v2, V.r2, nullptr, nullptr, st);  // v2 = A p2, (r2,v2)
  finish(spmv_parts<double>(A), A->partials, EPI_ALPHA, st);
  hipLaunchKernelGGL(k_cgs_q, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, V.u1, V.v2, V.q1,
                     V.w, st);
  double *pp1 = V.w;
  if (prec != 0) {
    prec_apply_S<double>(A, prec, V.w, V.p1, V.yL, V.zU, st);
    pp1 = V.p1;
  }
  // q2 = A p1 into the v2 slot (v2 is dead once q1 and w are formed)
  launch_spmv<double, SPMV_CGS_TAIL>(A, pp1, V.v2, nullptr, V.x, V.r, st);
  finish(spmv_parts<double>(A), A->partials, EPI_NORM, st);
}

// A persistent sweep that gave up waiting leaves a flag: report it (the
// results of that call are not trustworthy), clear it and re-arm the sweep
// outputs.
int check_sweep(tfs_matrix *A) {
  int e = 0;
  TFS_HIP(hipMemcpy(&e, A->sweep_err, sizeof(int), hipMemcpyDeviceToHost));
  if (!e) return 0;
  TFS_HIP(hipMemset(A->sweep_err, 0, sizeof(int)));
  if (A->vec64) arm(A->n, A->vec64 + 2 * (size_t)A->n);
  if (A->vec32) arm(A->n, A->vec32 + 2 * (size_t)A->n);
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  std::fprintf(stderr, "tfs: a preconditioner sweep timed out waiting for its dependencies\n");
  return TFS_ERR_DEVICE;
}

int read_state(tfs_matrix *A, int k = 0) {
  TFS_HIP(hipMemcpyAsync(A->h_st + k, A->d_st + k, sizeof(KState), hipMemcpyDeviceToHost,
                         cur_stream()));
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  if (g_tm.on) tm_harvest();
  return 0;
}

int init_state(tfs_matrix *A, int k, double tol, int cap, double ini_res) {
  KState *h = A->h_st + k;
  std::memset(h, 0, sizeof(KState));
  h->tol = tol;
  h->cap = cap;
  h->ini_res = ini_res;
  TFS_HIP(hipMemcpyAsync(A->d_st + k, h, sizeof(KState), hipMemcpyHostToDevice, cur_stream()));
  return 0;
}

// Queues iterations until the device state says done, reading the state
// between chunks.  The chunk is the number of iterations still needed at the
// observed contraction rate, so few are queued past convergence.
template <typename F>
int iterate(tfs_matrix *A, int cap, F &&one_iteration) {
  KState *h = A->h_st;
  if (!h->done && cap < 1) {  // zero-trip DO loop: iter = 1 (Cg.f90:97,181)
    h->done = 1;
    h->niter_out = 1;
    h->fin_res = h->err;
    h->status = kStNotConverged;
  }
  int queued = 0, chunk = 1, it_prev = h->iter;
  double err_prev = h->err;
  while (!h->done) {
    for (int k = 0; k < chunk; ++k) one_iteration();
    queued += chunk;
    TFS_HIP(hipGetLastError());
    TFS_TRY(read_state(A));
    if (h->done) break;
    if (queued > cap + 64) return TFS_ERR_DEVICE;  // the norm epilogue stops every solve at cap
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
  return 0;
}

int status_of(const KState *h) {
  if (h->status == kStBreakdown) return TFS_ERR_BREAKDOWN;
  return h->status == kStNotConverged ? TFS_NOT_CONVERGED : TFS_OK;
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
  TFS_TRY(ensure_vec64(A, method == TFS_CGS ? 12 : 7));
  TFS_TRY(prec_form<double>(A, prec));  // Cg.f90:42 (kept while the values are unchanged)
  const Vecs<double> V = vecs_of(A->vec64, n);
  KState *st = A->d_st;
  TFS_TRY(init_state(A, 0, tol, cap, *ini_res));
  const int g = vgrid(n);
  // caller order -> S order
  gather(n, x, A->s_orig, V.x);
  gather(n, r1, A->s_orig, V.r);
  // ||b|| and its early exit; r = b - A x; ||r|| and its early exit
  hipLaunchKernelGGL(k_sumsq<double>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, V.r,
                     (int)EPI_BNRM, st, A->partials);
  finish(g, A->partials, EPI_BNRM, st);
  launch_spmv<double, SPMV_RESID>(A, V.x, V.r, nullptr, nullptr, nullptr, st);
  if (method == TFS_CGS) gather(n, V.r, (const int *)nullptr, V.r2, st);  // r2 = r1 (Cgs.f90:75-77)
  hipLaunchKernelGGL(k_sumsq<double>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, V.r,
                     (int)EPI_INIRES, st, A->partials);
  finish(g, A->partials, EPI_INIRES, st);
  TFS_HIP(hipGetLastError());
  TFS_TRY(read_state(A));
  if (method == TFS_CGS)
    TFS_TRY(iterate(A, cap, [&] { cgs_iteration(A, prec, V, st); }));
  else
    TFS_TRY(iterate(A, cap, [&] { cg_iteration<double>(A, prec, V, st); }));
  // S order -> caller order (x and the final residual in r1)
  scatter(n, V.x, A->s_orig, x);
  scatter(n, V.r, A->s_orig, r1);
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  TFS_TRY(check_sweep(A));
  const KState *h = A->h_st;
  *niter = h->niter_out;
  *fin_res = h->fin_res;
  *ini_res = h->ini_res;  // still the caller's value after the ||b|| < tol exit
  return status_of(h);
}

// The sliced ELL of the reliable-update refinement's FP32 operator (see
// k_ell_diff), from the host layout of a matrix: the off-diagonal entries of
// each S row, 64 rows per slice.
int build_ell(tfs_matrix *A, const Layout &Ly) {
  const int n = A->n, ns = (n + kWave - 1) / kWave;
  std::vector<long long> off((size_t)ns + 1, 0);
  for (int sl = 0; sl < ns; ++sl) {
    int w = 0;
    for (int i = sl * kWave; i < std::min(n, (sl + 1) * kWave); ++i)
      w = std::max(w, Ly.s_row[i + 1] - Ly.s_row[i] - 1);
    off[(size_t)sl + 1] = off[(size_t)sl] + (long long)w * kWave;
  }
  const long long len = off[(size_t)ns];
  std::vector<int> ecol((size_t)len), esrc((size_t)len, -1);
  for (int sl = 0; sl < ns; ++sl) {
    const int w = (int)((off[(size_t)sl + 1] - off[(size_t)sl]) / kWave);
    for (int lane = 0; lane < kWave; ++lane) {
      const int i = sl * kWave + lane;
      int j = 0;
      if (i < n)
        for (int e = Ly.s_row[i]; e < Ly.s_row[i + 1]; ++e)
          if (Ly.s_col[e] != i) {
            const size_t q = (size_t)(off[(size_t)sl] + (long long)j * kWave + lane);
            ecol[q] = Ly.s_col[e];
            esrc[q] = Ly.s_src[e];
            ++j;
          }
      for (; j < w; ++j) ecol[(size_t)(off[(size_t)sl] + (long long)j * kWave + lane)] = i < n ? i : 0;
    }
  }
  A->ell_nslice = ns;
  A->ell_len = len;
  TFS_TRY(upload(A, &A->ell_off, off));
  TFS_TRY(upload(A, &A->ell_col, ecol));
  TFS_TRY(upload(A, &A->ell_src, esrc));
  TFS_TRY(dev_alloc(A, (void **)&A->ell_val, sizeof(float) * (size_t)std::max(len, 1LL)));
  TFS_TRY(dev_alloc(A, (void **)&A->ell_rsum, sizeof(float) * (size_t)std::max(n, 1)));
  A->ell_ready = false;
  return 0;
}

// The point-coloured sweep layout (k_colour_sweep) of a matrix whose rows
// are already in colour order: built only if its lower levels are exactly
// its colours in S order (true of a greedy point colouring, where a row's
// level is its colour), so that the L-order factor is also in S order.
int build_colour_sweeps(tfs_matrix *A, const Layout &Ly) {
  const int n = A->n;
  for (int p = 0; p < n; ++p)
    if (Ly.L.orig[p] != p || Ly.s_orig[p] != p) return 0;  // not a point colouring
  const int nc = Ly.L.nlev;
  std::vector<int> cstart(Ly.L.lev_ptr.begin(), Ly.L.lev_ptr.end()), sbeg(1, 0), row0;
  for (int c = 0; c < nc; ++c) {
    for (int r = cstart[c]; r < cstart[c + 1]; r += kWave) row0.push_back(r);
    sbeg.push_back((int)row0.size());
  }
  const int ns = (int)row0.size();
  for (int t = 0; t < 2; ++t) {  // 0: lower entries, 1: upper entries
    std::vector<long long> off((size_t)ns + 1, 0);
    auto take = [&](int i, int e) { return t == 0 ? Ly.s_col[e] < i : Ly.s_col[e] > i; };
    for (int c = 0; c < nc; ++c)
      for (int sl = sbeg[c]; sl < sbeg[c + 1]; ++sl) {
        int w = 0;
        for (int i = row0[sl]; i < std::min(row0[sl] + kWave, cstart[c + 1]); ++i) {
          int k = 0;
          for (int e = Ly.s_row[i]; e < Ly.s_row[i + 1]; ++e) k += take(i, e);
          w = std::max(w, k);
        }
        off[(size_t)sl + 1] = off[(size_t)sl] + (long long)w * kWave;
      }
    const long long len = off[(size_t)ns];
    std::vector<int> ecol((size_t)len, -1), esrc((size_t)len, -1);
    for (int c = 0; c < nc; ++c)
      for (int sl = sbeg[c]; sl < sbeg[c + 1]; ++sl)
        for (int lane = 0; lane < kWave; ++lane) {
          const int i = row0[sl] + lane;
          if (i >= cstart[c + 1]) break;
          int j = 0;
          for (int e = Ly.s_row[i]; e < Ly.s_row[i + 1]; ++e)
            if (take(i, e)) {
              const size_t q = (size_t)(off[(size_t)sl] + (long long)j * kWave + lane);
              ecol[q] = Ly.s_col[e];
              esrc[q] = Ly.s_src[e];
              ++j;
            }
        }
    A->csw_len[t] = len;
    TFS_TRY(upload(A, &A->csw_off[t], off));
    TFS_TRY(upload(A, &A->csw_col[t], ecol));
    TFS_TRY(upload(A, &A->csw_src[t], esrc));
    TFS_TRY(dev_alloc(A, (void **)&A->csw_val[t], sizeof(float) * (size_t)std::max(len, 1LL)));
  }
  TFS_TRY(upload(A, &A->csw_row0, row0));
  A->csw_cstart = cstart;
  A->csw_sbeg = sbeg;
  A->csw_colours = nc;
  A->csw_ready = false;
  return 0;
}

int prepare_colour_sweeps(tfs_matrix *A) {
  if (A->csw_ready || A->csw_colours == 0) return 0;
  for (int t = 0; t < 2; ++t)
    if (A->csw_len[t] > 0)
      hipLaunchKernelGGL(k_ell_values<float>, dim3(grid_for(A->csw_len[t], kRedThreads, 8192)),
                         dim3(kRedThreads), 0, cur_stream(), A->csw_len[t], A->csw_src[t], A->val,
                         A->csw_val[t]);
  TFS_HIP(hipGetLastError());
  A->csw_ready = true;
  return 0;
}

// z := M^-1 r for prec 2 on a point-coloured matrix (k_colour_sweep): the
// forward launches of colours 0..K-1 into y (the last writes z instead), then
// the backward launches of colours K-2..0 into z; all vectors in S order
// (FP32), factor d (its L order is its S order).  Each launch that writes z
// leaves per-workgroup partial sums of (r, z) at its own offset; returns
// their total count, for the rho reduction.
int colour_sweeps(tfs_matrix *A, const float *r, float *y, float *z, const KState *st,
                  bool y0_ready = false) {
  const float *d = A->v32.dL;
  const int K = A->csw_colours;
  // every launch that writes z keeps its own partial slots: at most
  // kMaxPartials over the K such launches
  static const int grid_cap = [] {
    const char *e = std::getenv("TFS_CSW_GRID");  // launch-shape experiments
    return e && std::atoi(e) > 0 ? std::atoi(e) : kSweepGrid;
  }();
  const int cap = std::max(1, std::min(grid_cap, kMaxPartials / std::max(1, K)));
  int parts = 0;
  auto launch = [&](int c, bool bwd, bool last) {
    const int s0 = A->csw_sbeg[c], s1 = A->csw_sbeg[c + 1], t = bwd ? 1 : 0;
    const int g = grid_for(std::max(1, s1 - s0), kEllWaves, cap);
    auto kern = bwd ? k_colour_sweep<true, false>
                    : last ? k_colour_sweep<false, true> : k_colour_sweep<false, false>;
    tm_begin(TFS_KC_SWEEP);
    hipLaunchKernelGGL(kern, dim3(g), dim3(kEllWaves * kWave), 0, cur_stream(), s0, s1,
                       A->csw_row0, A->csw_off[t], A->csw_col[t], A->csw_val[t], d, r, y, z,
                       A->csw_cstart[c + 1], st, A->partials + parts);
    tm_end(TFS_KC_SWEEP);
    if (bwd || last) parts += g;
  };
  // y0_ready: colour 0's forward step (y = r d, no lower entries) was done by
  // the kernel that last wrote r (k_axpy_mixed)
  for (int c = y0_ready && K >= 2 ? 1 : 0; c < K; ++c) launch(c, false, c == K - 1);
  for (int c = K - 2; c >= 0; --c) launch(c, true, false);
  return parts;
}

// The ELL values and row sums from the caller-order FP64 values.
int prepare_ell(tfs_matrix *A) {
  if (A->ell_ready) return 0;
  if (A->ell_len > 0)
    hipLaunchKernelGGL(k_ell_values<float>, dim3(grid_for(A->ell_len, kRedThreads, 8192)),
                       dim3(kRedThreads), 0, cur_stream(), A->ell_len, A->ell_src, A->val,
                       A->ell_val);
  if (A->n > 0)
    hipLaunchKernelGGL(k_row_sums, dim3(vgrid(A->n)), dim3(kRedThreads), 0, cur_stream(), A->n,
                       A->s_row, A->s_src, A->val, A->ell_rsum);
  TFS_HIP(hipGetLastError());
  A->ell_ready = true;
  return 0;
}

inline int ell_grid(const tfs_matrix *A) { return grid_for(A->ell_nslice, kEllWaves); }

// q = A32 p in difference form; DOT: partial sums of (p, q) per workgroup.
template <bool DOT>
void launch_ell(tfs_matrix *A, const float *x, float *y, const KState *st) {
  tm_begin(TFS_KC_SPMV32);
  hipLaunchKernelGGL(k_ell_diff<DOT>, dim3(ell_grid(A)), dim3(kEllWaves * kWave), 0, cur_stream(),
                     A->n, A->ell_nslice, A->ell_off, A->ell_col, A->ell_val, A->ell_rsum, x, y,
                     st, A->partials, kEllDiffXcd);
  tm_end(TFS_KC_SPMV32);
}

// The reordered copy used by the FP32 inner solves: P A P^T, P the block
// multi-colour order of this matrix's S-order graph (blocks of inner_block
// rows grown breadth-first, tfs_layout.hpp), or with inner_block 1 the greedy
// point colouring (red-black on a hex mesh), or the caller's own order
// (TFS_INNER_REFERENCE, used by the reliable-update loop).  Built from the device copy of
// the S layout (the caller's arrays are not kept).  The copy keeps its rows in
// that order for the Krylov vectors and the SpMV (identity S order); its L/U
// levels follow the order (one per colour for the point colouring, each
// colour's block-internal depth for blocks).
int build_inner(tfs_matrix *A) {
  if (A->inner) return 0;
  const int n = A->n;
  const long long nnz = A->nnz;
  std::vector<int> row(n + 1), col((size_t)nnz), src((size_t)nnz);
  TFS_HIP(hipMemcpyAsync(row.data(), A->s_row, sizeof(int) * (n + 1), hipMemcpyDeviceToHost,
                         cur_stream()));
  TFS_HIP(hipMemcpyAsync(col.data(), A->s_col, sizeof(int) * nnz, hipMemcpyDeviceToHost,
                         cur_stream()));
  TFS_HIP(hipMemcpyAsync(src.data(), A->s_src, sizeof(int) * nnz, hipMemcpyDeviceToHost,
                         cur_stream()));
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  std::vector<int> perm, bstart, bcolour;
  int ncol = 0;
  bool bricks = false;
  if (A->inner_order == TFS_INNER_REFERENCE) {
    // the caller's order (the reference factor, in FP32)
    perm.resize(n);
    TFS_HIP(hipMemcpy(perm.data(), A->s_of_orig, sizeof(int) * n, hipMemcpyDeviceToHost));
  } else if ((A->inner_order == TFS_INNER_BRICKS ||
              (A->inner_order == TFS_INNER_AUTO && A->inner_block == 1 && n >= kBrickAutoRows)) &&
             n > 0) {
    // bricks from the caller's numbering: its CRS rebuilt from the S layout
    std::vector<int> s_orig(n), s_of(n);
    TFS_HIP(hipMemcpy(s_orig.data(), A->s_orig, sizeof(int) * n, hipMemcpyDeviceToHost));
    TFS_HIP(hipMemcpy(s_of.data(), A->s_of_orig, sizeof(int) * n, hipMemcpyDeviceToHost));
    std::vector<int> crow(n + 1, 0), ccol((size_t)nnz), cdia(n, 0);
    for (int i = 0; i < n; ++i) {
      const int k = s_of[i];
      crow[i + 1] = crow[i] + row[k + 1] - row[k];
      int q = crow[i];
      for (int e = row[k]; e < row[k + 1]; ++e, ++q) {
        ccol[(size_t)q] = s_orig[col[e]];  // the S row keeps the caller's entry order
        if (ccol[(size_t)q] == i) cdia[i] = q;
      }
    }
    std::vector<int> pc;
    const int edge = brick_edge(n);
    A->inner_brick = edge;
    ncol = brick_colour_order(n, crow.data(), ccol.data(), cdia.data(), edge, edge * edge * edge, pc,
                              bstart, bcolour);
    perm.resize(n);
    for (int i = 0; i < n; ++i) perm[i] = s_of[pc[i]];
    bricks = true;
    if (A->inner_order == TFS_INNER_AUTO && ncol != 2) {  // irregular blocks: points
      bricks = false;
      ncol = colour_order(n, row.data(), col.data(), perm, nullptr, 1);
    }
  } else {
    ncol = A->inner_block > 1
               ? block_colour_order(n, row.data(), col.data(), A->inner_block, perm)
               : colour_order(n, row.data(), col.data(), perm, nullptr, 1);
  }
  std::vector<int> pos(n);
  for (int i = 0; i < n; ++i) pos[perm[i]] = i;
  if (bricks || (A->inner_order != TFS_INNER_REFERENCE && A->inner_block > 1)) {
    // Rows of a block order (swept by the generic sweeps) sorted by their
    // forward level, ties in block order.  Every coupling keeps its direction
    // (a row's earlier neighbours sit on lower levels), so the factor is the
    // same; the copy's S order is then its L order and the U order its
    // reverse, so the forward sweep reads its right-hand side and k_rho the
    // backward sweep's output in contiguous runs instead of gathering them
    // row by row from many blocks.  The FP32 SpMV's gathers lose some
    // locality in this order.  Measured per solve (MI355X, Poisson box of
    // aspect 100): 16^3 bricks at 256^3 2094 -> 1972 ms (forward sweep
    // 304 -> 252 us, k_rho 73 -> 55, SpMV 207 -> 221), 8^3 at 192^3 734 -> 723
    // ms, 4^3 at 256^3 2510 -> 2526 ms.
    std::vector<int> lev(n, 0), cnt;
    for (int i = 0; i < n; ++i) {
      const int sr = perm[i];
      int l = 0;
      for (int k = row[sr]; k < row[sr + 1]; ++k) {
        const int j = pos[col[k]];
        if (j < i) l = std::max(l, lev[j] + 1);
      }
      lev[i] = l;
      if (l + 1 > (int)cnt.size()) cnt.resize(l + 1, 0);
      cnt[l]++;
    }
    std::vector<int> start(cnt.size() + 1, 0), sorted(n);
    for (size_t l = 0; l < cnt.size(); ++l) start[l + 1] = start[l] + cnt[l];
    for (int i = 0; i < n; ++i) sorted[start[lev[i]]++] = perm[i];
    perm.swap(sorted);
    for (int i = 0; i < n; ++i) pos[perm[i]] = i;
  }

  std::vector<int> prow(n + 1, 0), pcol((size_t)nnz), pdia(n, -1), emap((size_t)nnz);
  std::vector<std::pair<int, int>> ent;
  long long q = 0;
  for (int i = 0; i < n; ++i) {
    const int s = perm[i];
    ent.clear();
    for (int k = row[s]; k < row[s + 1]; ++k) ent.emplace_back(pos[col[k]], src[k]);
    std::sort(ent.begin(), ent.end());
    for (const auto &e : ent) {
      if (e.first == i) pdia[i] = (int)q;
      pcol[(size_t)q] = e.first;
      emap[(size_t)q] = e.second;
      ++q;
    }
    prow[i + 1] = (int)q;
  }
  Layout Ly;
  tfs_matrix *C = nullptr;
  TFS_TRY(matrix_create(n, nnz, prow.data(), pcol.data(), pdia.data(), 0, &C, &Ly, true));
  std::vector<int> from_s(n);
  for (int i = 0; i < n; ++i) from_s[i] = perm[Ly.s_orig[i]];
  int rc = upload(A, &A->inner_emap, emap);
  if (rc == 0) rc = upload(A, &A->inner_from_s, from_s);
  if (rc == 0) rc = build_ell(C, Ly);
  // the point-coloured copy sweeps with k_colour_sweep; the brick-ordered one
  // with the synchronisation-free sweeps (its DAG is a few levels deep: 4E - 2
  // for E^3 bricks in two colours, 14 / 30 / 62 for E = 4 / 8 / 16)
  if (rc == 0 && !bricks && A->inner_order != TFS_INNER_REFERENCE && A->inner_block == 1)
    rc = build_colour_sweeps(C, Ly);
  if (rc < 0) {
    tfs_matrix_destroy(C);
    return rc;
  }
  C->persistent = A->persistent;
  C->dynamic = A->dynamic;
  C->coop = A->coop;
  if (A->grid_fixed) {  // a caller's limit holds for the copy too
    C->grid_persistent = std::min(C->grid_persistent, A->grid_persistent);
    C->grid_fixed = true;
  }
  A->inner = C;
  A->inner_colours = ncol;
  A->inner_built = A->inner_order == TFS_INNER_REFERENCE ? TFS_INNER_REFERENCE
                   : bricks                              ? TFS_INNER_BRICKS
                                                         : TFS_INNER_COLOURED;
  A->inner_ready = false;
  A->bytes += C->bytes;
  return 0;
}

// The inner copy's values from this matrix's (caller-order gather through
// the entry map), marking the copy's layouts stale.
int refresh_inner(tfs_matrix *A) {
  tfs_matrix *I = A->inner;
  if (A->inner_ready) return 0;
  gather((int)A->nnz, A->val, A->inner_emap, I->val);
  values_changed(I);
  A->inner_ready = true;
  TFS_HIP(hipGetLastError());
  return 0;
}

// One iteration of the reliable-update refinement's FP32 CG on the inner copy
// I (Cg.f90:97-178 with FP32 vectors): z = M^-1 r (FP32 factor and sweeps),
// rho = (r, z), p = z + beta p, q = A32 p in difference form with (p, q),
// c += alpha p into the FP64 correction, r -= alpha q with ||r||.
void reliable_iteration(tfs_matrix *I, int prec, const Vecs<float> &F, double *c, KState *st,
                        bool y0_ready) {
  const int n = I->n, g = vgrid(n);
  const bool csw = prec == 2 && I->csw_colours > 0;
  float *z = prec == 0 ? F.r : F.z;
  if (csw) {
    // z in S order (forward result in slot r2), with the rho partial sums
    finish(colour_sweeps(I, F.r, F.r2, F.z, st, y0_ready), I->partials, EPI_RHO, st);
  }
  // z stays in the backward sweep's (reversed) order when that is the S
  // order backwards: k_rho and the p update read it from there
  const bool zrev = prec == 2 && !csw && I->s_to_u_rev;
  if (!csw) {
    if (prec == 2) dic_apply<float>(I, F.r, F.yL, F.zU, st);
    const int mode = prec == 2 ? (zrev ? 5 : 2) : prec == 1 ? 1 : 0;
    hipLaunchKernelGGL(k_rho<float>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, F.r, F.z,
                       vals<float>(I).s_diag, F.zU, I->s_to_u, mode, st, I->partials);
    finish(g, I->partials, EPI_RHO, st);
  }
  if (zrev)
    hipLaunchKernelGGL(k_pupdate_rev<float>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n,
                       F.zU, F.p, st);
  else
    hipLaunchKernelGGL(k_pupdate<float>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, z, F.p,
                       st);
  launch_ell<true>(I, F.p, F.q, st);
  finish(ell_grid(I), I->partials, EPI_ALPHA, st);
  // with point-coloured sweeps the tail also does the next sweep's colour-0
  // step (y0_ready in the next iteration)
  const int n0 = csw && I->csw_colours >= 2 ? I->csw_cstart[1] : 0;
  hipLaunchKernelGGL(k_axpy_mixed, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, c, F.r, F.p,
                     F.q, st, I->partials, n0, I->v32.dL, F.r2);
  finish(g, I->partials, EPI_NORM, st);
}

int push_state(tfs_matrix *A, int k) {
  TFS_HIP(hipMemcpyAsync(A->d_st + k, A->h_st + k, sizeof(KState), hipMemcpyHostToDevice,
                         cur_stream()));
  return 0;
}

// FP32 CG with reliable FP64 updates (TFS_REFINE_RELIABLE; after the reliable
// updates of van der Vorst & Ye, "Residual replacement strategies for Krylov
// subspace iterative methods for the convergence of true residuals", SIAM J.
// Sci. Comput. 22 (2000), as used for mixed-precision CG by Clark et al.,
// Comput. Phys. Commun. 181 (2010)).  One CG recurrence runs on the inner copy
// in FP32 -- factor, sweeps, SpMV (difference form) and the vectors r, z, p,
// q -- while its directions are folded every iteration into an FP64
// correction c.  Whenever the FP32 recurrence residual has fallen by `delta`
// against the last true residual, the segment ends: x += c (FP64), r = b - A x
// with the FP64 matrix, and the recurrence continues from that true residual
// with its direction p and rho kept (no restart, so no Krylov information is
// lost).  It stops when the true residual is below tol.  niter counts the
// residual replacements, inner_iters the CG iterations.
int solve_mixed_device(tfs_matrix *A, int prec, double *x, double *r1, int *niter,
                       int *inner_iters, double tol, double inner_rtol, double x_rtol,
                       double *ini_res, double *fin_res);

// FP32 iterations one segment of the reliable-update loop may take to reduce
// its residual by delta; a segment that needs more ends the solve as not
// converged (as FP64 CG would fail on such a system, e.g. an unpreconditioned
// nonsymmetric one).
constexpr int kSegmentCap = 20000;

int solve_reliable_device(tfs_matrix *A, int prec, double *x, double *r1, int *niter,
                          int *inner_iters, double tol, double delta, double x_rtol,
                          double *ini_res, double *fin_res) {
  const int n = A->n, cap = *niter;
  if (!(delta > 0.0)) delta = 0.1;
  delta = std::min(0.5, std::max(1e-6, delta));
  TFS_TRY(build_inner(A));
  tfs_matrix *I = A->inner;
  TFS_TRY(refresh_inner(A));
  const int *to_s = A->inner_from_s;  // I's S position -> A's S position
  TFS_TRY(ensure_vec64(A, 8));
  TFS_TRY(ensure_vec32(I, 8));
  TFS_TRY(prepare_values<double>(A));
  TFS_TRY(prec_form<float>(I, prec));
  TFS_TRY(prepare_ell(I));
  if (prec == 2) TFS_TRY(prepare_colour_sweeps(I));
  const Vecs<double> D = vecs_of(A->vec64, n);
  double *bS = D.r2, *c = D.z;  // c: the FP64 correction, in I's S order
  const Vecs<float> F = vecs_of(I->vec32, n);
  KState *si = I->d_st;
  KState *ho = A->h_st + 1, *hi = I->h_st;
  const int g = vgrid(n);
  gather(n, x, A->s_orig, D.x);
  gather(n, r1, A->s_orig, bS);
  auto true_residual = [&]() -> int {
    TFS_TRY(init_state(A, 1, tol, cap, 0.0));
    gather(n, bS, (const int *)nullptr, D.r);
    launch_spmv<double, SPMV_RESID>(A, D.x, D.r, nullptr, nullptr, nullptr, nullptr);
    hipLaunchKernelGGL(k_sumsq<double>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, D.r,
                       (int)EPI_BNRM, A->d_st + 1, A->partials);
    finish(g, A->partials, EPI_BNRM, A->d_st + 1);
    TFS_HIP(hipGetLastError());
    return read_state(A, 1);
  };
  TFS_TRY(true_residual());
  const double err0 = ho->bnrm2;
  double err = err0, best = err0;
  hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, c, 0.0);
  int updates = 0, stalls = 0, status = TFS_OK;
  bool started = false;
  while (err >= tol) {
    if (updates >= cap) {
      status = TFS_NOT_CONVERGED;
      break;
    }
    // the recurrence continues from the true residual, in FP32
    gather(n, D.r, to_s, F.r);
    const double seg_tol = std::max(delta * err, 0.5 * tol);
    if (!started) {
      TFS_TRY(init_state(I, 0, seg_tol, kSegmentCap, err0));
      hi->err = err;
    } else {
      hi->done = 0;
      hi->status = 0;
      hi->tol = seg_tol;
      hi->err = err;
      hi->cap = hi->iter + kSegmentCap;
      hi->rho_old = hi->rho;  // the last iteration's rho (kept on its exit)
    }
    TFS_TRY(push_state(I, 0));
    started = true;
    // the first iteration of a segment starts from a replaced residual: its
    // sweep does colour 0's step itself; later ones find it done by the tail
    bool seg_first = true;
    TFS_TRY(iterate(I, hi->cap, [&] {
      reliable_iteration(I, prec, F, c, si, !seg_first);
      seg_first = false;
    }));
    if (hi->status == kStBreakdown) {
      status = TFS_ERR_BREAKDOWN;
      break;
    }
    // residual replacement: x += c; r = b - A x in FP64
    ++updates;
    hipLaunchKernelGGL(k_fold_correction, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, D.x, c,
                       to_s);
    TFS_TRY(true_residual());
    const double e_new = ho->bnrm2;
    if (!std::isfinite(e_new)) {
      err = e_new;
      status = TFS_ERR_BREAKDOWN;
      break;
    }
    err = e_new;
    if (hi->status == kStNotConverged) {
      if (err >= tol) status = TFS_NOT_CONVERGED;
      break;
    }
    // no progress of the true residual over three segments: FP32 can no
    // longer reduce it
    stalls = e_new >= best ? stalls + 1 : 0;
    best = std::min(best, e_new);
    if (stalls >= 3) {
      if (err >= tol) status = TFS_NOT_CONVERGED;
      break;
    }
  }
  *inner_iters = started ? hi->iter : 0;
  scatter(n, D.x, A->s_orig, x);
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  TFS_TRY(check_sweep(A));
  TFS_TRY(check_sweep(I));
  *ini_res = err0;
  if (status == TFS_OK && updates > 0 && x_rtol > 0.0 && updates < cap) {
    // The FP32 recurrence cannot see the part of the residual below FP32
    // resolution (error along the smallest eigenvectors).  For the solution
    // criterion the flexible-CG loop, whose directions are FP64, continues
    // from x (b back in r1) until its last update is below x_rtol ||x||.
    scatter(n, bS, A->s_orig, r1);
    TFS_HIP(hipGetLastError());
    int steps = cap - updates, inner = 0;
    double ir = 0.0, fr = 0.0;
    A->refine = TFS_REFINE_FCG;
    const int rc = solve_mixed_device(A, prec, x, r1, &steps, &inner, tol, 0.0, x_rtol, &ir, &fr);
    A->refine = TFS_REFINE_RELIABLE;
    *inner_iters += inner;
    *fin_res = fr;
    *niter = rc == TFS_NOT_CONVERGED ? cap + 1 : updates + steps;
    return rc;
  }
  scatter(n, D.r, A->s_orig, r1);
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  *fin_res = err;
  *niter = status == TFS_NOT_CONVERGED ? cap + 1 : updates;
  return status;
}

// FP32-inner PCG inside an FP64 outer loop (iterative refinement).
//   The outer loop is flexible CG (Notay's FCG(1)) in FP64: r = b - A x;
//   z = B(r), the FP32 PCG solve of A z = r (values, factor and vectors in
//   FP32) to ||r_in|| < inner_rtol ||r||; p = z + beta p with
//   beta = -(z, A p_prev) / (p_prev, A p_prev); alpha = (p, r) / (p, A p);
//   x += alpha p; r -= alpha A p (FP64 SpMV) -- until ||r|| < tol, checked on
//   the true residual b - A x.  With TFS_REFINE_DEFECT the outer loop is plain
//   defect correction (x += z).  Defect correction contracts the error by
//   about cond(A) * 2^-24 per step (the FP32 rounding of A), which stalls near
//   cond(A) ~ 1e7; the Krylov outer loop only needs the inner solve to be a
//   good preconditioner, whatever the conditioning.
//   x_rtol > 0: the last update must also satisfy ||dx|| <= x_rtol ||x||.
// The outer residual, updates and every dot product stay in FP64 on the device.
int solve_mixed_device(tfs_matrix *A, int prec, double *x, double *r1, int *niter,
                       int *inner_iters, double tol, double inner_rtol, double x_rtol,
                       double *ini_res, double *fin_res) {
  if (!A || !niter || !inner_iters || !ini_res || !fin_res) return TFS_ERR_INVALID;
  if (prec < 0 || prec > 2) prec = 0;
  const int n = A->n, cap = *niter;
  *inner_iters = 0;
  if (n == 0) {
    *niter = 0;
    *fin_res = 0.0;
    return TFS_OK;
  }
  if (!x || !r1) return TFS_ERR_INVALID;
  if (A->refine == TFS_REFINE_RELIABLE)
    return solve_reliable_device(A, prec, x, r1, niter, inner_iters, tol, inner_rtol, x_rtol,
                                 ini_res, fin_res);
  if (!(inner_rtol > 0.0)) inner_rtol = 1e-2;
  inner_rtol = std::min(0.5, std::max(1e-6, inner_rtol));
  // The inner solves run on I: this matrix, or its colour-ordered copy when
  // the factor's order matters (prec 2).  The diagonal preconditioners do not
  // depend on the order.
  const bool coloured = A->inner_order != TFS_INNER_REFERENCE && prec == 2;
  const bool krylov = A->refine == TFS_REFINE_FCG;
  tfs_matrix *I = A;
  if (coloured) {
    TFS_TRY(build_inner(A));
    I = A->inner;
    TFS_TRY(refresh_inner(A));
  }
  const int *to_s = coloured ? A->inner_from_s : nullptr;  // I's S position -> A's
  TFS_TRY(ensure_vec64(A, 8));
  TFS_TRY(ensure_vec32(I, 7));
  TFS_TRY(prepare_values<double>(A));
  TFS_TRY(prec_form<float>(I, prec));
  // FP64: x, r (residual), z (inner result), p, q = A p, r2 slot = b.  The
  // sweep slots (yL, zU) are not touched: they stay armed for FP64 solves.
  const Vecs<double> D = vecs_of(A->vec64, n);
  double *bS = D.r2;
  const Vecs<float> F = vecs_of(I->vec32, n);    // inner solve: x = correction
  KState *so = A->d_st + 1, *si = I->d_st;
  KState *ho = A->h_st + 1;
  const int g = vgrid(n);
  gather(n, x, A->s_orig, D.x);
  gather(n, r1, A->s_orig, bS);
  auto dot = [&](const double *u, const double *v, int epi) {
    hipLaunchKernelGGL(k_dot<double>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, u, v,
                       A->partials);
    finish(g, A->partials, epi, so);
  };
  auto sumsq = [&](const double *u, int epi) {
    hipLaunchKernelGGL(k_sumsq<double>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, u, epi,
                       so, A->partials);
    finish(g, A->partials, epi, so);
  };
  // true residual r = b - A x and its norm (FP64) into bnrm2; with
  // update_norms also ||z|| and ||x|| (defect correction's update criterion)
  auto true_residual = [&](bool update_norms = false) -> int {
    TFS_TRY(init_state(A, 1, tol, cap, 0.0));
    if (update_norms) {
      hipLaunchKernelGGL(k_sumsq<float>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, F.x,
                         (int)EPI_NAUX4, so, A->partials);
      finish(g, A->partials, EPI_NAUX4, so);
      sumsq(D.x, EPI_NAUX5);
    }
    gather(n, bS, (const int *)nullptr, D.r);
    launch_spmv<double, SPMV_RESID>(A, D.x, D.r, nullptr, nullptr, nullptr, nullptr);
    sumsq(D.r, EPI_BNRM);
    TFS_HIP(hipGetLastError());
    return read_state(A, 1);
  };
  TFS_TRY(true_residual());
  const double err0 = ho->bnrm2;
  double err = err0;
  bool small_update = x_rtol <= 0.0;  // no criterion on the update
  bool have_p = false;                // a previous direction to orthogonalise against
  double pq_prev = 0.0;
  int outer = 0, status = TFS_OK, stalls = 0;
  while (err >= tol || !small_update) {
    if (outer >= cap) {
      status = TFS_NOT_CONVERGED;
      break;
    }
    ++outer;
    // inner solve, FP32: e = 0, r32 = r (in I's S order)
    const double inner_tol =
        krylov || err < tol ? inner_rtol * err : std::max(inner_rtol * err, 0.5 * tol);
    gather(n, D.r, to_s, F.r);
    hipLaunchKernelGGL(k_fill<float>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, F.x, 0.0f);
    TFS_TRY(init_state(I, 0, inner_tol, 100000, 0.0));
    hipLaunchKernelGGL(k_sumsq<float>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, F.r,
                       (int)EPI_INIRES, si, I->partials);
    finish(g, I->partials, EPI_INIRES, si);
    TFS_HIP(hipGetLastError());
    TFS_TRY(read_state(I));
    TFS_TRY(iterate(I, 100000, [&] { cg_iteration<float>(I, prec, F, si); }));
    *inner_iters += std::min(I->h_st->niter_out, 100000);
    if (I->h_st->status == kStBreakdown) {
      status = TFS_ERR_BREAKDOWN;
      break;
    }
    double e_new = 0.0;
    if (!krylov) {
      // defect correction: x += z; r = b - A x
      hipLaunchKernelGGL(k_add_correction, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, D.x,
                         F.x, to_s);
      TFS_TRY(true_residual(!small_update));
      e_new = ho->bnrm2;
      if (!small_update) small_update = ho->aux[4] <= x_rtol * ho->aux[5];
    } else {
      // z (FP64, this matrix's S order) from the inner result
      hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, D.z, 0.0);
      hipLaunchKernelGGL(k_add_correction, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, D.z,
                         F.x, to_s);
      if (have_p) {
        dot(D.z, D.q, EPI_AUX0);  // (z, A p_prev)
        TFS_HIP(hipGetLastError());
        TFS_TRY(read_state(A, 1));
        const double beta = -ho->aux[0] / pq_prev;
        hipLaunchKernelGGL(k_xpby, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, D.z, D.p, beta);
      } else {
        gather(n, D.z, (const int *)nullptr, D.p);
      }
      launch_spmv<double, SPMV_DOT>(A, D.p, D.q, D.p, nullptr, nullptr, nullptr);  // q = A p
      finish(spmv_parts<double>(A), A->partials, EPI_AUX1, so);                   // (p, q)
      dot(D.p, D.r, EPI_AUX2);                                                     // (p, r)
      if (!small_update) {
        sumsq(D.p, EPI_NAUX4);
        sumsq(D.x, EPI_NAUX5);
      }
      TFS_HIP(hipGetLastError());
      TFS_TRY(read_state(A, 1));
      const double pq = ho->aux[1], alpha = ho->aux[2] / pq;
      if (!(pq > 0.0) || !std::isfinite(alpha)) {
        status = TFS_ERR_BREAKDOWN;
        break;
      }
      if (!small_update) small_update = std::fabs(alpha) * ho->aux[4] <= x_rtol * ho->aux[5];
      hipLaunchKernelGGL(k_update_xr, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, D.x, D.r,
                         D.p, D.q, alpha, A->partials);
      finish(g, A->partials, EPI_NAUX3, so);
      TFS_HIP(hipGetLastError());
      TFS_TRY(read_state(A, 1));
      e_new = ho->aux[3];
      pq_prev = pq;
      have_p = true;
      if (e_new < tol) {
        // confirm on the true residual; if the recurrence drifted, restart from it
        TFS_TRY(true_residual());
        e_new = ho->bnrm2;
        if (e_new >= tol) have_p = false;
      }
    }
    if (!std::isfinite(e_new)) {
      err = e_new;
      status = TFS_ERR_BREAKDOWN;
      break;
    }
    // no progress: defect correction less than 10 % per step, the Krylov
    // loop (whose residual need not fall monotonically) no fall at all
    stalls = e_new > (krylov ? 1.0 : 0.9) * err ? stalls + 1 : 0;
    err = e_new;
    if (stalls >= 3) {  // FP32 can no longer reduce the defect
      if (err >= tol) status = TFS_NOT_CONVERGED;
      break;             // (below tol: the update is at the rounding floor)
    }
  }
  if (krylov && status == TFS_OK) {
    // the returned residual is the true one (the recurrence's may differ)
    TFS_TRY(true_residual());
    err = ho->bnrm2;
  }
  scatter(n, D.x, A->s_orig, x);
  scatter(n, D.r, A->s_orig, r1);
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  TFS_TRY(check_sweep(A));
  if (I != A) TFS_TRY(check_sweep(I));
  *ini_res = err0;
  *fin_res = err;
  *niter = status == TFS_NOT_CONVERGED ? cap + 1 : outer;
  return status;
}

// ---- SpMV shape: rows per row block (one wavefront each) and wavefronts per
// workgroup.  Row sums do not depend on it (each row is summed left to right
// by one lane); the dot products fused into the SpMV are summed per
// workgroup, so their rounding follows the shape.

// Row blocks over the S rows: at most `rows` rows and kSpmvCap entries each
// (a longer row is a block of its own).
std::vector<int> row_blocks(const std::vector<int> &s_row, int rows) {
  const int n = (int)s_row.size() - 1;
  std::vector<int> rb(1, 0);
  for (int r = 0; r < n;) {
    int e = r + 1;
    while (e < n && e - r < rows && s_row[e + 1] - s_row[r] <= kSpmvCap) ++e;
    rb.push_back(e);
    r = e;
  }
  return rb;
}

int download_s_row(tfs_matrix *A, std::vector<int> *s_row) {
  s_row->resize((size_t)A->n + 1);
  TFS_HIP(hipMemcpyAsync(s_row->data(), A->s_row, sizeof(int) * (A->n + 1),
                         hipMemcpyDeviceToHost, cur_stream()));
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  return 0;
}

int set_spmv_shape(tfs_matrix *A, const std::vector<int> &s_row, int rows, int waves) {
  const std::vector<int> rb = row_blocks(s_row, rows);
  int *d = nullptr;
  TFS_HIP(hipStreamSynchronize(cur_stream()));  // no launch still reads the old blocks
  TFS_TRY(upload(A, &d, rb));
  TFS_TRY(dev_free(A, A->rblk, sizeof(int) * ((size_t)A->nblk + 1)));
  A->rblk = d;
  A->nblk = (int)rb.size() - 1;
  A->spmv_rows = rows;
  A->spmv_waves = waves;
  return 0;
}

// Times the plain FP64 SpMV for rows per block 16/32/64 x wavefronts per
// workgroup 2/4/8 (20 back-to-back launches each) and keeps the fastest.
int tune_spmv(tfs_matrix *A) {
  if (A->n == 0) return 0;
  std::vector<int> s_row;
  TFS_TRY(download_s_row(A, &s_row));
  const bool was_ready = A->v64.ready;
  TFS_TRY(prepare_values<double>(A));
  double *xy = nullptr;
  const size_t nb = sizeof(double) * (size_t)A->n * 2;
  TFS_TRY(dev_alloc(A, (void **)&xy, nb));
  hipLaunchKernelGGL(k_fill<double>, dim3(vgrid(A->n)), dim3(kRedThreads), 0, cur_stream(), A->n,
                     xy, 1.0);
  float best = 0.f;
  int br = A->spmv_rows, bw = A->spmv_waves;
  for (int rows : {16, 32, 64}) {
    TFS_TRY(set_spmv_shape(A, s_row, rows, 4));
    for (int waves : {2, 4, 8}) {
      A->spmv_waves = waves;
      launch_spmv<double, SPMV_PLAIN>(A, xy, xy + A->n, nullptr, nullptr, nullptr, nullptr);
      TFS_HIP(hipEventRecord(A->ev0, cur_stream()));
      for (int k = 0; k < 20; ++k)
        launch_spmv<double, SPMV_PLAIN>(A, xy, xy + A->n, nullptr, nullptr, nullptr, nullptr);
      TFS_HIP(hipEventRecord(A->ev1, cur_stream()));
      TFS_HIP(hipGetLastError());
      TFS_HIP(hipEventSynchronize(A->ev1));
      float t = 0.f;
      TFS_HIP(hipEventElapsedTime(&t, A->ev0, A->ev1));
      if (best == 0.f || t < best) {
        best = t;
        br = rows;
        bw = waves;
      }
    }
  }
  TFS_TRY(set_spmv_shape(A, s_row, br, bw));
  TFS_TRY(dev_free(A, xy, nb));
  if (!was_ready) values_changed(A);
  return 0;
}

int upload_sweep(tfs_matrix *A, const SweepLayout &S, DevSweep &D) {
  D.nlev = S.nlev;
  D.lev_ptr = S.lev_ptr;
  D.segs = S.segs;
  D.ntri = (long long)S.col.size();
  TFS_TRY(upload(A, &D.lev, S.lev_ptr));
  TFS_TRY(upload(A, &D.ptr, S.ptr));
  TFS_TRY(upload(A, &D.col, S.col));
  TFS_TRY(upload(A, &D.of_orig, S.of_orig));
  return 0;
}

// The FP64 sliced ELL of all entries (k_ell64) from the host layout: 64 S
// rows per slice, each row's entries in the caller's order, column -1 as
// padding.  Not built when the padding would exceed a quarter of nnz (rows
// of very uneven length); the CSR kernel serves then.
int build_ell64(tfs_matrix *A, const Layout &Ly) {
  const int n = A->n, ns = (n + kWave - 1) / kWave;
  std::vector<long long> off((size_t)ns + 1, 0);
  for (int sl = 0; sl < ns; ++sl) {
    int w = 0;
    for (int i = sl * kWave; i < std::min(n, (sl + 1) * kWave); ++i)
      w = std::max(w, Ly.s_row[i + 1] - Ly.s_row[i]);
    off[(size_t)sl + 1] = off[(size_t)sl] + (long long)w * kWave;
  }
  const long long len = off[(size_t)ns];
  if (n == 0 || (double)len > 1.25 * (double)A->nnz) return 0;
  std::vector<int> ecol((size_t)len, -1), esrc((size_t)len, -1);
  for (int i = 0; i < n; ++i) {
    const int sl = i / kWave, lane = i % kWave;
    int j = 0;
    for (int e = Ly.s_row[i]; e < Ly.s_row[i + 1]; ++e, ++j) {
      const size_t q = (size_t)(off[(size_t)sl] + (long long)j * kWave + lane);
      ecol[q] = Ly.s_col[e];
      esrc[q] = Ly.s_src[e];
    }
  }
  TFS_TRY(upload(A, &A->e64_off, off));
  TFS_TRY(upload(A, &A->e64_col, ecol));
  TFS_TRY(upload(A, &A->e64_src, esrc));
  TFS_TRY(dev_alloc(A, (void **)&A->e64_val, sizeof(double) * (size_t)len));
  A->e64_nslice = ns;
  A->e64_len = len;
  return 0;
}

template <typename T, int OP>
int occupancy(int *per_cu) {
  int b = 0;
  TFS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &b, reinterpret_cast<const void *>(k_sweep_syncfree<T, OP, false>), kChunk, 0));
  *per_cu = std::min(*per_cu, b);
  return 0;
}

int create_device(tfs_matrix *A, const Layout &Ly) {
  const int n = A->n;
  TFS_HIP(hipStreamCreateWithFlags(&A->stream, hipStreamNonBlocking));
  TFS_HIP(hipEventCreate(&A->ev0));
  TFS_HIP(hipEventCreate(&A->ev1));
  StreamScope scope(A->stream);
  TFS_TRY(upload(A, &A->s_row, Ly.s_row));
  TFS_TRY(upload(A, &A->s_col, Ly.s_col));
  TFS_TRY(upload(A, &A->s_src, Ly.s_src));
  TFS_TRY(upload(A, &A->s_diag_src, Ly.s_diag_src));
  TFS_TRY(upload(A, &A->s_orig, Ly.s_orig));
  TFS_TRY(upload(A, &A->s_of_orig, Ly.s_of_orig));
  TFS_TRY(upload(A, &A->rblk, Ly.s_rblk));
  A->nblk = (int)Ly.s_rblk.size() - 1;
  TFS_TRY(build_ell64(A, Ly));
  TFS_TRY(upload_sweep(A, Ly.L, A->L));
  TFS_TRY(upload_sweep(A, Ly.U, A->U));
  TFS_TRY(upload(A, &A->l_to_s, Ly.l_to_s));
  TFS_TRY(upload(A, &A->u_to_l, Ly.u_to_l));
  TFS_TRY(upload(A, &A->s_to_u, Ly.s_to_u));
  A->l_ident = A->u_rev = A->s_to_u_rev = true;
  for (int i = 0; i < n; ++i) {
    A->l_ident = A->l_ident && Ly.l_to_s[(size_t)i] == i;
    A->u_rev = A->u_rev && Ly.u_to_l[(size_t)i] == n - 1 - i;
    A->s_to_u_rev = A->s_to_u_rev && Ly.s_to_u[(size_t)i] == n - 1 - i;
  }
  {
    std::vector<int> lpos((size_t)n);
    for (int k = 0; k < n; ++k) lpos[(size_t)k] = Ly.L.of_orig[(size_t)Ly.s_orig[(size_t)k]];
    TFS_TRY(upload(A, &A->s_lpos, lpos));
  }
  const size_t nb = sizeof(double) * (size_t)std::max(n, 1);
  TFS_TRY(dev_alloc(A, (void **)&A->val, sizeof(double) * (size_t)std::max(A->nnz, 1LL)));
  TFS_TRY(dev_alloc(A, (void **)&A->stage_x, nb));
  TFS_TRY(dev_alloc(A, (void **)&A->stage_r, nb));
  TFS_TRY(dev_alloc(A, (void **)&A->d_st, 2 * sizeof(KState)));
  TFS_TRY(dev_alloc(A, (void **)&A->partials, sizeof(double) * kMaxPartials));
  TFS_TRY(dev_alloc(A, (void **)&A->sweep_err, sizeof(int)));
  TFS_HIP(hipMemset(A->sweep_err, 0, sizeof(int)));
  TFS_TRY(dev_alloc(A, (void **)&A->sweep_ctr, 2 * sizeof(unsigned int)));
  TFS_HIP(hipMemset(A->sweep_ctr, 0, 2 * sizeof(unsigned int)));
  TFS_HIP(hipMemset(A->d_st, 0, 2 * sizeof(KState)));
  TFS_HIP(hipMemset(A->val, 0, sizeof(double) * (size_t)std::max(A->nnz, 1LL)));
  TFS_HIP(hipHostMalloc((void **)&A->h_st, 2 * sizeof(KState), hipHostMallocDefault));
  // the FP64 value layouts every FP64 path uses, allocated here so that the
  // allocation-free entry point finds them
  TFS_TRY(alloc_values<double>(A));
  // Persistent sweeps: at most as many workgroups as can be co-resident for
  // the heaviest sweep kernel (the static schedule needs them all resident).
  int dev = 0, n_cu = 0, per_cu = 8;
  TFS_HIP(hipGetDevice(&dev));
  TFS_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  TFS_TRY((occupancy<double, OP_FACTOR>(&per_cu)));
  TFS_TRY((occupancy<double, OP_FWD>(&per_cu)));
  TFS_TRY((occupancy<double, OP_BWD>(&per_cu)));
  TFS_TRY((occupancy<float, OP_FACTOR>(&per_cu)));
  TFS_TRY((occupancy<float, OP_FWD>(&per_cu)));
  TFS_TRY((occupancy<float, OP_BWD>(&per_cu)));
  // At most kSweepGroups workgroups: every waiting row polls memory, and with
  // the whole chip's worth of pollers the polls delay the stores they wait
  // for (measured on MI355X, 19-level momentum 128^3 sweep: 73 us with 768
  // workgroups, 75 with 1024, 85 with 1280, 100 with 1536 or more; the
  // 766-level Poisson 256^3 one: 1.15 ms with 768, 1.18 with 1024).
  const int chunks = (n + kChunk - 1) / kChunk;
  A->sweep_capacity = std::max(1, n_cu * std::max(1, per_cu));
  A->grid_persistent = std::max(1, std::min({chunks, A->sweep_capacity, kSweepGroups}));
  A->persistent = per_cu > 0;
  if (const char *e = std::getenv("TFS_SWEEP")) {
    if (std::strcmp(e, "levels") == 0) A->persistent = false;
    if (std::strcmp(e, "dynamic") == 0) A->dynamic = true;
    if (std::strcmp(e, "cooperative") == 0) A->coop = true;
    A->sweep_fixed = true;  // "static" (or anything else): the persistent static schedule
  }
  if (const char *e = std::getenv("TFS_SWEEP_GROUPS")) {  // the cap, up to the occupancy
    const int g = std::atoi(e);
    if (g > 0) {
      A->grid_persistent = std::max(1, std::min({chunks, n_cu * std::max(1, per_cu), g}));
      A->grid_fixed = true;
    }
  }
  A->bandwidth = Ly.bandwidth;
  // the uploads and fills above ran on the null stream; the matrix stream is
  // non-blocking, so finish them before any work is queued there (the null
  // stream only: other matrices' streams keep running)
  TFS_HIP(hipStreamSynchronize(nullptr));
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
    case TFS_CANCELLED: return "cancelled before it ran";
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

namespace {

// Validates and rebases the structure (Matrix_Mod.f90:15-34 layout) before
// touching the device, so a bad structure is rejected without a GPU; then
// builds the layouts and uploads them.  `keep` receives the host layout.
int matrix_create(int n, long long nnz, const int *row, const int *col, const int *dia, int base,
                  tfs_matrix **out, Layout *keep, bool identity_s) {
  if (!out) return TFS_ERR_INVALID;
  *out = nullptr;
  if (n < 0 || nnz < n || nnz > 0x7ffffffeLL || (base != 0 && base != 1)) return TFS_ERR_INVALID;
  if (n > 0 && (!row || !col || !dia)) return TFS_ERR_INVALID;
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

  Layout local;
  Layout &Ly = keep ? *keep : local;
  build_layout(n, hr.data(), hc.data(), hd.data(), kSpmvRows, kSpmvCap, kSmallLevel, Ly,
               identity_s);
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

}  // namespace

extern "C" int tfs_matrix_create(int n, long long nnz, const int *row, const int *col,
                                 const int *dia, int base, tfs_matrix **out) {
  return matrix_create(n, nnz, row, col, dia, base, out, nullptr);
}

extern "C" int tfs_matrix_destroy(tfs_matrix *A) {
  if (!A) return TFS_OK;
  if (A->inner) tfs_matrix_destroy(A->inner);
  if (A->stream) (void)hipStreamSynchronize(A->stream);
  for (void *p : A->allocs) (void)hipFree(p);
  if (A->h_st) (void)hipHostFree(A->h_st);
  if (A->ev0) (void)hipEventDestroy(A->ev0);
  if (A->ev1) (void)hipEventDestroy(A->ev1);
  if (A->stream) (void)hipStreamDestroy(A->stream);
  delete A;
  return TFS_OK;
}

extern "C" int tfs_matrix_set_values(tfs_matrix *A, const double *val) {
  if (!A) return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  if (!A || (A->nnz > 0 && !val)) return TFS_ERR_INVALID;
  TFS_HIP(hipMemcpyAsync(A->val, val, sizeof(double) * (size_t)A->nnz, hipMemcpyHostToDevice,
                         cur_stream()));
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  values_changed(A);
  return TFS_OK;
}

extern "C" int tfs_matrix_set_values_device(tfs_matrix *A, const double *d_val) {
  if (!A) return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  if (!A || (A->nnz > 0 && !d_val)) return TFS_ERR_INVALID;
  TFS_HIP(hipMemcpyAsync(A->val, d_val, sizeof(double) * (size_t)A->nnz, hipMemcpyDeviceToDevice,
                         cur_stream()));
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  values_changed(A);
  return TFS_OK;
}

extern "C" int tfs_matrix_get_values(tfs_matrix *A, double *val) {
  if (!A) return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  if (A->nnz > 0 && !val) return TFS_ERR_INVALID;
  TFS_HIP(hipMemcpyAsync(val, A->val, sizeof(double) * (size_t)A->nnz, hipMemcpyDeviceToHost,
                         cur_stream()));
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  return TFS_OK;
}

extern "C" int tfs_matrix_values_changed(tfs_matrix *A) {
  if (!A) return TFS_ERR_INVALID;
  values_changed(A);
  return TFS_OK;
}

extern "C" double *tfs_matrix_values_device(tfs_matrix *A) { return A ? A->val : nullptr; }

hipStream_t tfs::matrix_stream(const tfs_matrix *A) { return A ? A->stream : nullptr; }
int tfs::sweep_capacity(const tfs_matrix *A) { return A ? A->sweep_capacity : 1; }
void tfs::set_sweep_grid(tfs_matrix *A, int groups) {
  if (!A) return;
  A->grid_persistent = std::max(1, std::min(A->grid_persistent, groups));
  A->grid_fixed = true;
  if (A->inner) set_sweep_grid(A->inner, groups);
}

extern "C" int tfs_matrix_get_info(const tfs_matrix *A, tfs_matrix_info *info) {
  if (!A || !info) return TFS_ERR_INVALID;
  info->n = A->n;
  info->nnz = A->nnz;
  info->levels_lower = A->L.nlev;
  info->levels_upper = A->U.nlev;
  info->row_blocks = A->nblk;
  info->device_bytes = A->bytes;
  info->inner_colours = A->inner_colours;
  info->spmv_rows = A->spmv_rows;
  info->spmv_waves = A->spmv_waves;
  info->inner_block = A->inner_block;
  info->inner_levels = A->inner ? A->inner->L.nlev : 0;
  info->ell64_slices = A->e64_len > 0 ? A->e64_nslice : 0;
  info->inner_order = A->inner ? A->inner_built : -1;
  info->inner_brick_edge = A->inner && A->inner_built == TFS_INNER_BRICKS ? A->inner_brick : 0;
  info->sweep_groups = A->persistent ? A->grid_persistent : 0;
  info->inner_sweep_groups =
      A->inner ? (A->inner->persistent ? A->inner->grid_persistent : 0) : -1;
  return TFS_OK;
}

extern "C" int tfs_matrix_tune_spmv(tfs_matrix *A, int *rows_per_block, int *waves_per_group) {
  if (!A) return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  TFS_TRY(tune_spmv(A));
  if (rows_per_block) *rows_per_block = A->spmv_rows;
  if (waves_per_group) *waves_per_group = A->spmv_waves;
  return TFS_OK;
}

extern "C" int tfs_matrix_set_spmv_shape(tfs_matrix *A, int rows_per_block, int waves_per_group) {
  if (!A || rows_per_block < 1 || rows_per_block > kSpmvRows ||
      (waves_per_group != 2 && waves_per_group != 4 && waves_per_group != 8))
    return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  if (A->n == 0) return TFS_OK;
  std::vector<int> s_row;
  TFS_TRY(download_s_row(A, &s_row));
  return set_spmv_shape(A, s_row, rows_per_block, waves_per_group);
}

namespace {
// Drops the inner copy (rebuilt on the next refinement solve).
int drop_inner(tfs_matrix *A) {
  if (!A->inner) return 0;
  StreamScope scope(A->stream);
  TFS_HIP(hipStreamSynchronize(A->stream));
  A->bytes -= A->inner->bytes;
  tfs_matrix_destroy(A->inner);
  A->inner = nullptr;
  TFS_TRY(dev_free(A, A->inner_emap, sizeof(int) * (size_t)A->nnz));
  TFS_TRY(dev_free(A, A->inner_from_s, sizeof(int) * (size_t)A->n));
  A->inner_emap = A->inner_from_s = nullptr;
  A->inner_colours = 0;
  return 0;
}
}  // namespace

extern "C" int tfs_matrix_set_inner_block(tfs_matrix *A, int rows) {
  if (!A || rows < 1 || rows > 4096) return TFS_ERR_INVALID;
  if (rows == A->inner_block) return TFS_OK;
  A->inner_block = rows;
  return drop_inner(A);
}

extern "C" int tfs_matrix_set_refinement(tfs_matrix *A, int outer) {
  if (!A || (outer != TFS_REFINE_FCG && outer != TFS_REFINE_DEFECT && outer != TFS_REFINE_RELIABLE))
    return TFS_ERR_INVALID;
  A->refine = outer;
  return TFS_OK;
}

extern "C" int tfs_matrix_set_inner_order(tfs_matrix *A, int order) {
  if (!A || order < TFS_INNER_REFERENCE || order > TFS_INNER_AUTO) return TFS_ERR_INVALID;
  if (order == A->inner_order) return TFS_OK;
  A->inner_order = order;
  return drop_inner(A);
}

extern "C" int tfs_solve_device(tfs_matrix *A, int method, int prec, double *d_x, double *d_r1,
                                int *niter, double tol, double *ini_res, double *fin_res) {
  if (!A) return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  return solve_device(A, method, prec, d_x, d_r1, niter, tol, ini_res, fin_res);
}

extern "C" int tfs_solve(tfs_matrix *A, int method, int prec, double *x1, double *r1, int *niter,
                         double tol, double *ini_res, double *fin_res) {
  if (!A) return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  if (A->n > 0 && (!x1 || !r1)) return TFS_ERR_INVALID;
  const size_t bytes = sizeof(double) * (size_t)A->n;
  if (A->n > 0) {
    TFS_HIP(hipMemcpyAsync(A->stage_x, x1, bytes, hipMemcpyHostToDevice, cur_stream()));
    TFS_HIP(hipMemcpyAsync(A->stage_r, r1, bytes, hipMemcpyHostToDevice, cur_stream()));
  }
  const int rc = solve_device(A, method, prec, A->stage_x, A->stage_r, niter, tol, ini_res, fin_res);
  if (rc < 0 && rc != TFS_ERR_BREAKDOWN) return rc;
  if (A->n > 0) {
    TFS_HIP(hipMemcpyAsync(x1, A->stage_x, bytes, hipMemcpyDeviceToHost, cur_stream()));
    TFS_HIP(hipMemcpyAsync(r1, A->stage_r, bytes, hipMemcpyDeviceToHost, cur_stream()));
    TFS_HIP(hipStreamSynchronize(cur_stream()));
  }
  return rc;
}

extern "C" int tfs_solve_mixed_device(tfs_matrix *A, int prec, double *d_x, double *d_r1,
                                      int *niter, int *inner_iters, double tol,
                                      double inner_rtol, double x_rtol, double *ini_res,
                                      double *fin_res) {
  if (!A) return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  return solve_mixed_device(A, prec, d_x, d_r1, niter, inner_iters, tol, inner_rtol, x_rtol,
                            ini_res, fin_res);
}

// Building blocks on caller-order vectors: permuted into the S layout and back.
extern "C" int tfs_spmv_device(tfs_matrix *A, const double *d_x, double *d_y) {
  if (!A) return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  if (!A || (A->n > 0 && (!d_x || !d_y))) return TFS_ERR_INVALID;
  if (A->n == 0) return TFS_OK;
  TFS_TRY(ensure_vec64(A, 7));
  TFS_TRY(prepare_values<double>(A));
  const Vecs<double> V = vecs_of(A->vec64, A->n);
  gather(A->n, d_x, A->s_orig, V.x);
  launch_spmv<double, SPMV_PLAIN>(A, V.x, V.q, nullptr, nullptr, nullptr, nullptr);
  scatter(A->n, V.q, A->s_orig, d_y);
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  return TFS_OK;
}

extern "C" int tfs_residual_device(tfs_matrix *A, const double *d_x, double *d_r) {
  if (!A) return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  if (!A || (A->n > 0 && (!d_x || !d_r))) return TFS_ERR_INVALID;
  if (A->n == 0) return TFS_OK;
  TFS_TRY(ensure_vec64(A, 7));
  TFS_TRY(prepare_values<double>(A));
  const Vecs<double> V = vecs_of(A->vec64, A->n);
  gather(A->n, d_x, A->s_orig, V.x);
  gather(A->n, d_r, A->s_orig, V.r);
  launch_spmv<double, SPMV_RESID>(A, V.x, V.r, nullptr, nullptr, nullptr, nullptr);
  scatter(A->n, V.r, A->s_orig, d_r);
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  return TFS_OK;
}

// tfs_residual_device without allocating: the two S-order vectors live in
// this device's `scratch` array (the FP64 value layouts exist from creation;
// the values are copied into them if they changed, as in the variant that
// allocates).  Every scratch element the call reads it has written first, so
// nothing needs clearing; calls that overlap on one GPU share the array and
// race.
extern "C" int tfs_residual_device_noalloc(tfs_matrix *A, const double *d_x, double *d_r) {
  if (!A) return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  if (A->n > 0 && (!d_x || !d_r)) return TFS_ERR_INVALID;
  if (A->n == 0) return TFS_OK;
  if (A->n > kScratchRows) {
    std::fprintf(stderr,
                 "tfs_residual_device_noalloc: %d rows exceed the scratch capacity of %d rows\n",
                 A->n, kScratchRows);
    return TFS_ERR_INVALID;
  }
  double *base = g_lib.scratch[g_lib.device];
  if (!base || !A->v64.allocated) return TFS_ERR_DEVICE;
  TFS_TRY(prepare_values<double>(A));
  double *xS = base, *rS = base + kScratchRows;
  gather(A->n, d_x, A->s_orig, xS);
  gather(A->n, d_r, A->s_orig, rS);
  launch_spmv<double, SPMV_RESID>(A, xS, rS, nullptr, nullptr, nullptr, nullptr);
  scatter(A->n, rS, A->s_orig, d_r);
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  return TFS_OK;
}

extern "C" int tfs_prec_form(tfs_matrix *A, int prec) {
  if (!A) return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  if (!A) return TFS_ERR_INVALID;
  TFS_TRY(prec_form<double>(A, prec));
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  TFS_TRY(check_sweep(A));
  return TFS_OK;
}

extern "C" int tfs_prec_solve_device(tfs_matrix *A, int prec, const double *d_b, double *d_x) {
  if (!A) return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  if (!A || (A->n > 0 && (!d_b || !d_x))) return TFS_ERR_INVALID;
  if (prec < 0 || prec > 2) prec = 0;
  if (A->v64.factor_prec != prec) return TFS_ERR_INVALID;  // tfs_prec_form first
  if (A->n == 0) return TFS_OK;
  TFS_TRY(ensure_vec64(A, 7));
  const Vecs<double> V = vecs_of(A->vec64, A->n);
  gather(A->n, d_b, A->s_orig, V.r);
  prec_apply_S<double>(A, prec, V.r, V.z, V.yL, V.zU, nullptr);
  scatter(A->n, V.z, A->s_orig, d_x);
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  TFS_TRY(check_sweep(A));
  return TFS_OK;
}

extern "C" int tfs_prec_diag_device(tfs_matrix *A, double *d_out) {
  if (!A) return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  if (!A || (A->n > 0 && !d_out)) return TFS_ERR_INVALID;
  const int fp = A->v64.factor_prec;
  if (fp < 0) return TFS_ERR_INVALID;
  if (A->n == 0) return TFS_OK;
  if (fp == 2) {
    gather(A->n, A->v64.dL, A->L.of_orig, d_out);
  } else if (fp == 1) {
    gather(A->n, A->v64.s_diag, A->s_of_orig, d_out);
  } else {
    hipLaunchKernelGGL(k_fill<double>, dim3(vgrid(A->n)), dim3(kRedThreads), 0, cur_stream(),
                       A->n, d_out, 1.0);
  }
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  return TFS_OK;
}

extern "C" int tfs_prec_diag32_device(tfs_matrix *A, int prec, double *d_out) {
  if (!A) return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  if (!A || (A->n > 0 && !d_out)) return TFS_ERR_INVALID;
  if (prec < 0 || prec > 2) prec = 0;
  if (A->n == 0) return TFS_OK;
  TFS_TRY(prec_form<float>(A, prec));
  if (prec == 2) {
    gather(A->n, A->v32.dL, A->L.of_orig, d_out);
  } else if (prec == 1) {
    gather(A->n, A->v32.s_diag, A->s_of_orig, d_out);
  } else {
    hipLaunchKernelGGL(k_fill<double>, dim3(vgrid(A->n)), dim3(kRedThreads), 0, cur_stream(),
                       A->n, d_out, 1.0);
  }
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  TFS_TRY(check_sweep(A));
  return TFS_OK;
}

extern "C" int tfs_spmv_time(tfs_matrix *A, const double *d_x, double *d_y, int reps, double *ms) {
  if (!A) return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  if (!A || !ms || reps < 1 || (A->n > 0 && (!d_x || !d_y))) return TFS_ERR_INVALID;
  if (A->n == 0) {
    *ms = 0.0;
    return TFS_OK;
  }
  TFS_TRY(prepare_values<double>(A));
  // The vectors are taken as S-ordered: the kernel's traffic is the same.
  launch_spmv<double, SPMV_PLAIN>(A, d_x, d_y, nullptr, nullptr, nullptr, nullptr);  // warm
  TFS_HIP(hipEventRecord(A->ev0, cur_stream()));
  for (int k = 0; k < reps; ++k)
    launch_spmv<double, SPMV_PLAIN>(A, d_x, d_y, nullptr, nullptr, nullptr, nullptr);
  TFS_HIP(hipEventRecord(A->ev1, cur_stream()));
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipEventSynchronize(A->ev1));
  float t = 0.f;
  TFS_HIP(hipEventElapsedTime(&t, A->ev0, A->ev1));
  *ms = (double)t / reps;
  return TFS_OK;
}

namespace {

// Reads m words and stores nothing unless their xor is a value the caller
// never places there: a read-only sweep over a buffer larger than the
// memory-side cache leaves it holding clean lines of that buffer only, so a
// kernel timed next finds none of its own data cached and has no dirty
// lines to write back.
__global__ void __launch_bounds__(kRedThreads)
k_flush_read(long long m, const unsigned long long *__restrict__ buf, unsigned long long *out) {
  unsigned long long acc = 0;
  for (long long i = blockIdx.x * (long long)kRedThreads + threadIdx.x; i < m;
       i += (long long)gridDim.x * kRedThreads)
    acc ^= buf[i];
  if (acc == 0x5a5a5a5a5a5a5a5aULL) *out = acc;
}

}  // namespace

extern "C" int tfs_spmv_time_cold(tfs_matrix *A, const double *d_x, double *d_y, int reps,
                                  double *ms) {
  if (!A) return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  if (!ms || reps < 1 || (A->n > 0 && (!d_x || !d_y))) return TFS_ERR_INVALID;
  if (A->n == 0) {
    *ms = 0.0;
    return TFS_OK;
  }
  TFS_TRY(prepare_values<double>(A));
  // 640 MiB read before every launch: more than the 256 MB memory-side
  // cache, so the SpMV finds none of its data there
  const long long words = (640LL << 20) / 8;
  unsigned long long *flush = nullptr;
  TFS_HIP(hipMalloc((void **)&flush, (size_t)words * 8 + 8));
  float total = 0.f;
  int rc = 0;
  for (int k = 0; k < reps && rc == 0; ++k) {
    hipLaunchKernelGGL(k_flush_read, dim3(4096), dim3(kRedThreads), 0, cur_stream(), words, flush,
                       flush + words);
    if (hipGetLastError() != hipSuccess || hipEventRecord(A->ev0, cur_stream()) != hipSuccess) {
      rc = TFS_ERR_DEVICE;
      break;
    }
    launch_spmv<double, SPMV_PLAIN>(A, d_x, d_y, nullptr, nullptr, nullptr, nullptr);
    float t = 0.f;
    if (hipEventRecord(A->ev1, cur_stream()) != hipSuccess ||
        hipEventSynchronize(A->ev1) != hipSuccess ||
        hipEventElapsedTime(&t, A->ev0, A->ev1) != hipSuccess)
      rc = TFS_ERR_DEVICE;
    total += t;
  }
  (void)hipGetLastError();
  (void)hipFree(flush);
  if (rc) return rc;
  *ms = (double)total / reps;
  return TFS_OK;
}

extern "C" int tfs_sweep_time(tfs_matrix *A, int reps, double *ms) {
  if (!A) return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  if (!ms || reps < 1) return TFS_ERR_INVALID;
  if (A->n == 0) {
    *ms = 0.0;
    return TFS_OK;
  }
  TFS_TRY(ensure_vec64(A, 7));
  TFS_TRY(prec_form<double>(A, 2));
  const Vecs<double> V = vecs_of(A->vec64, A->n);
  hipLaunchKernelGGL(k_fill<double>, dim3(vgrid(A->n)), dim3(kRedThreads), 0, cur_stream(), A->n,
                     V.z, 1.0);
  dic_apply<double>(A, V.z, V.yL, V.zU, nullptr);  // warm
  TFS_HIP(hipEventRecord(A->ev0, cur_stream()));
  for (int k = 0; k < reps; ++k) dic_apply<double>(A, V.z, V.yL, V.zU, nullptr);
  TFS_HIP(hipEventRecord(A->ev1, cur_stream()));
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipEventSynchronize(A->ev1));
  float t = 0.f;
  TFS_HIP(hipEventElapsedTime(&t, A->ev0, A->ev1));
  TFS_TRY(check_sweep(A));
  *ms = (double)t / (2.0 * reps);
  return TFS_OK;
}

extern "C" int tfs_timing_start(int kclass) {
  if (kclass < TFS_KC_SPMV || kclass > TFS_KC_SPMV32) return TFS_ERR_INVALID;
  TFS_TRY(ensure_init());
  TFS_HIP(hipDeviceSynchronize());
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
  TFS_HIP(hipDeviceSynchronize());
  tm_harvest();
  g_tm.on = false;
  if (ms) *ms = g_tm.ms;
  if (launches) *launches = g_tm.launches;
  // the events go back now: none outlives the measurement into process exit
  for (hipEvent_t e : g_tm.ev) (void)hipEventDestroy(e);
  g_tm.ev.clear();
  return TFS_OK;
}

// ============================================================================
// Batched solves: many independent systems of uneven size in one
// block-diagonal matrix.  The layouts, SpMV and sweeps run unchanged on the
// combined matrix (blocks never couple, so the combined DIC factor is the
// per-system factors); vector work and reductions go per system through
// tiles, and each system has its own device state.  One iteration of all
// systems is captured once into a hipGraph and replayed.

struct tfs_batch {
  int nsys = 0;
  tfs_matrix *A = nullptr;
  std::vector<int> n_of, row_off;
  std::vector<long long> nnz_of, nnz_off;
  int ntiles = 0;
  int *tile_sys = nullptr, *tile_beg = nullptr, *tile_end = nullptr, *tile_first = nullptr;
  // states: [0, nsys) inner, [nsys, 2 nsys) outer, 2 nsys inner-all, 2 nsys + 1 outer-all
  KState *d_st = nullptr;
  KState *h_st = nullptr;
  double *partials = nullptr;
  // [0] FP64 iteration, [1] FP32 (flexible-CG inner solves), [2] FP32 reliable updates
  hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};
  int graph_prec[3] = {-1, -1, -1};
  bool use_graph = true;
  int graph_launches = 0;
  // FP32 inner solves of the refinement on a colour-ordered copy (each
  // system's rows grouped by colour, systems in order), built on first use
  int inner_order = TFS_INNER_BRICKS;  // the refinement's copy: bricks per system
  int inner_block = kInnerBlock;
  int refine = TFS_REFINE_RELIABLE;  // outer loop of the refinement
  tfs_batch *inner = nullptr;
  int *inner_emap = nullptr;    // inner's entry k <- entry of A->val
  int *inner_from_s = nullptr;  // S position of A for inner's S position i
};

namespace {

constexpr int kTileRows = 2048;
constexpr int kGraphIters = 4;  // iterations per captured graph

int batch_alloc(tfs_batch *B, void **p, size_t bytes) { return dev_alloc(B->A, p, bytes); }

Tiles tiles_of(const tfs_batch *B) { return Tiles{B->tile_sys, B->tile_beg, B->tile_end}; }

void finish_seg(tfs_batch *B, int epi, KState *st, KState *all) {
  hipLaunchKernelGGL(k_finish_seg, dim3(B->nsys), dim3(kRedThreads), 0, cur_stream(),
                     B->tile_first, B->partials, epi, st, all);
}

// One CG iteration of every system (Cg.f90:97-178); converged systems idle.
template <typename T>
void batch_cg_iteration(tfs_batch *B, int prec, const Vecs<T> &V, KState *st, KState *all) {
  tfs_matrix *A = B->A;
  const Tiles tl = tiles_of(B);
  const int nt = B->ntiles;
  if (prec == 2) dic_apply<T>(A, V.r, V.yL, V.zU, all);
  const int mode = prec == 2 ? 2 : prec == 1 ? 1 : 0;
  T *z = prec == 0 ? V.r : V.z;
  hipLaunchKernelGGL(k_rho_seg<T>, dim3(nt), dim3(kRedThreads), 0, cur_stream(), tl, V.r, V.z,
                     vals<T>(A).s_diag, V.zU, A->s_to_u, mode, st, all, B->partials);
  finish_seg(B, EPI_RHO, st, all);
  hipLaunchKernelGGL(k_pupdate_seg<T>, dim3(nt), dim3(kRedThreads), 0, cur_stream(), tl, z, V.p,
                     st, all);
  launch_spmv<T, SPMV_PLAIN>(A, V.p, V.q, nullptr, nullptr, nullptr, all);
  hipLaunchKernelGGL(k_dot_seg<T>, dim3(nt), dim3(kRedThreads), 0, cur_stream(), tl, V.p, V.q, st,
                     all, B->partials);
  finish_seg(B, EPI_ALPHA, st, all);
  hipLaunchKernelGGL(k_axpy_norm_seg<T>, dim3(nt), dim3(kRedThreads), 0, cur_stream(), tl, V.x,
                     V.r, V.p, V.q, st, all, B->partials);
  finish_seg(B, EPI_NORM, st, all);
}

// Captures kGraphIters batched iterations (`one()`) into the graph of `slot`
// (once per slot and preconditioner).  If the capture fails -- e.g. under the
// cooperative schedule, whose launch is not capturable -- it is retried with
// per-level sweeps, and without a graph as a last resort.
template <typename F>
int batch_graph_fn(tfs_batch *B, int slot, int prec, F &&one, hipGraphExec_t *exec) {
  *exec = nullptr;
  if (!B->use_graph) return 0;
  if (B->graph[slot] && B->graph_prec[slot] == prec) {
    *exec = B->graph[slot];
    return 0;
  }
  if (B->graph[slot]) {
    TFS_HIP(hipGraphExecDestroy(B->graph[slot]));
    B->graph[slot] = nullptr;
  }
  for (int attempt = 0; attempt < 2; ++attempt) {
    const bool was_persistent = B->A->persistent;
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(cur_stream(), hipStreamCaptureModeRelaxed) != hipSuccess) break;
    for (int k = 0; k < kGraphIters; ++k) one();
    const hipError_t e1 = hipStreamEndCapture(cur_stream(), &g);
    hipError_t e2 = hipErrorUnknown;
    if (e1 == hipSuccess && g) e2 = hipGraphInstantiate(&B->graph[slot], g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    if (e1 == hipSuccess && e2 == hipSuccess && B->A->persistent == was_persistent) {
      B->graph_prec[slot] = prec;
      *exec = B->graph[slot];
      return 0;
    }
    if (B->graph[slot]) {
      (void)hipGraphExecDestroy(B->graph[slot]);
      B->graph[slot] = nullptr;
    }
    B->A->persistent = false;  // retry with per-level sweeps
  }
  std::fprintf(stderr, "tfs: batch iteration not capturable; launching kernels directly\n");
  B->use_graph = false;
  return 0;
}

template <typename T>
int batch_graph(tfs_batch *B, int prec, const Vecs<T> &V, KState *st, KState *all,
                hipGraphExec_t *exec) {
  return batch_graph_fn(B, sizeof(T) == 8 ? 0 : 1, prec,
                        [&] { batch_cg_iteration<T>(B, prec, V, st, all); }, exec);
}

int batch_read(tfs_batch *B) {
  TFS_HIP(hipMemcpyAsync(B->h_st, B->d_st, sizeof(KState) * (2 * B->nsys + 2),
                         hipMemcpyDeviceToHost, cur_stream()));
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  return 0;
}

int batch_write(tfs_batch *B) {
  TFS_HIP(hipMemcpyAsync(B->d_st, B->h_st, sizeof(KState) * (2 * B->nsys + 2),
                         hipMemcpyHostToDevice, cur_stream()));
  return 0;
}

// Replays iterations (`one()`, captured in the graph of `slot`) until every
// system is done (states read between chunks).
template <typename F>
int batch_iterate_fn(tfs_batch *B, int slot, int prec, KState *st, KState *all, int cap,
                     F &&one) {
  hipGraphExec_t exec = nullptr;
  TFS_TRY(batch_graph_fn(B, slot, prec, one, &exec));
  const int step = exec ? kGraphIters : 1;
  const KState *ha = B->h_st + (all - B->d_st);
  int queued = 0, chunk = step;
  std::vector<int> it_prev(B->nsys, 0);
  std::vector<double> err_prev(B->nsys, 0.0);
  const KState *hs = B->h_st + (st - B->d_st);
  for (int s = 0; s < B->nsys; ++s) {
    err_prev[s] = hs[s].err;
    it_prev[s] = hs[s].iter;
  }
  while (!ha->done) {
    for (int k = 0; k < chunk; k += step) {
      if (exec) {
        TFS_HIP(hipGraphLaunch(exec, cur_stream()));
        B->graph_launches++;
      } else {
        one();
      }
    }
    queued += chunk;
    TFS_HIP(hipGetLastError());
    TFS_TRY(batch_read(B));
    if (ha->done) break;
    if (queued > cap + 64 + step) return TFS_ERR_DEVICE;
    // the slowest unfinished system sets the next chunk
    int next = 1;
    for (int s = 0; s < B->nsys; ++s) {
      if (hs[s].done) continue;
      int need = 16;
      const int dit = hs[s].iter - it_prev[s];
      if (dit > 0 && hs[s].err > 0 && err_prev[s] > 0 && hs[s].err < err_prev[s]) {
        const double rate = std::pow(hs[s].err / err_prev[s], 1.0 / dit);
        const double it_need = std::log(hs[s].tol / hs[s].err) / std::log(rate);
        need = it_need < 1.0 ? 1 : (int)(0.9 * it_need);
      }
      next = std::max(next, need);
      it_prev[s] = hs[s].iter;
      err_prev[s] = hs[s].err;
    }
    chunk = std::max(step, std::min(32, next) / step * step);
  }
  return 0;
}

template <typename T>
int batch_iterate(tfs_batch *B, int prec, const Vecs<T> &V, KState *st, KState *all, int cap) {
  return batch_iterate_fn(B, sizeof(T) == 8 ? 0 : 1, prec, st, all, cap,
                          [&] { batch_cg_iteration<T>(B, prec, V, st, all); });
}

// caller per-system vectors <-> combined S-order vector
int batch_in(tfs_batch *B, double *const *src, double *stage, double *dstS) {
  for (int s = 0; s < B->nsys; ++s)
    if (B->n_of[s])
      TFS_HIP(hipMemcpyAsync(stage + B->row_off[s], src[s], sizeof(double) * B->n_of[s],
                             hipMemcpyDeviceToDevice, cur_stream()));
  gather(B->A->n, stage, B->A->s_orig, dstS);
  return 0;
}
int batch_out(tfs_batch *B, const double *srcS, double *stage, double *const *dst) {
  scatter(B->A->n, srcS, B->A->s_orig, stage);
  for (int s = 0; s < B->nsys; ++s)
    if (B->n_of[s])
      TFS_HIP(hipMemcpyAsync(dst[s], stage + B->row_off[s], sizeof(double) * B->n_of[s],
                             hipMemcpyDeviceToDevice, cur_stream()));
  return 0;
}

void batch_init_states(tfs_batch *B, int base, const int *cap, const double *tol) {
  for (int s = 0; s < B->nsys; ++s) {
    KState &h = B->h_st[base + s];
    std::memset(&h, 0, sizeof(KState));
    h.tol = tol[s];
    h.cap = cap[s];
  }
  KState &all = B->h_st[2 * B->nsys + (base ? 1 : 0)];
  std::memset(&all, 0, sizeof(KState));
  all.cap = B->nsys;
}

}  // namespace

namespace {

// The device side of a batch from its concatenated 0-based block-diagonal
// structure (B->nsys, n_of, row_off, nnz_of, nnz_off set): matrix, tiles,
// states, SpMV shape.  `Ly` receives the host layout.  On failure the
// caller destroys B.
int batch_build(tfs_batch *B, long long N, long long NNZ, const int *cr, const int *cc,
                const int *cd, Layout &Ly) {
  int rc = matrix_create((int)N, NNZ, cr, cc, cd, 0, &B->A, &Ly);
  if (rc < 0) return rc;
  B->A->sweep_fixed = true;  // the batch picks its sweeps when it captures its graph
  const int nsys = B->nsys;
  // every system occupies one contiguous S range (a connected block is laid
  // out contiguously by RCM; several components of one system stay adjacent)
  std::vector<int> tsys, tbeg, tend, tfirst;
  for (int s = 0; s < nsys; ++s) {
    int lo = (int)N, hi = -1;
    for (int i = 0; i < B->n_of[s]; ++i) {
      const int k = Ly.s_of_orig[B->row_off[s] + i];
      lo = std::min(lo, k);
      hi = std::max(hi, k);
    }
    if (hi - lo + 1 != B->n_of[s]) return TFS_ERR_INVALID;
    tfirst.push_back((int)tsys.size());
    for (int b = lo; b <= hi; b += kTileRows) {
      tsys.push_back(s);
      tbeg.push_back(b);
      tend.push_back(std::min(hi + 1, b + kTileRows));
    }
  }
  tfirst.push_back((int)tsys.size());
  B->ntiles = (int)tsys.size();
  tfs_matrix *A = B->A;
  if ((rc = upload(A, &B->tile_sys, tsys)) < 0 || (rc = upload(A, &B->tile_beg, tbeg)) < 0 ||
      (rc = upload(A, &B->tile_end, tend)) < 0 || (rc = upload(A, &B->tile_first, tfirst)) < 0 ||
      (rc = batch_alloc(B, (void **)&B->d_st, sizeof(KState) * (2 * nsys + 2))) < 0 ||
      (rc = batch_alloc(B, (void **)&B->partials, sizeof(double) * std::max(1, B->ntiles))) < 0)
    return rc;
  if (hipHostMalloc((void **)&B->h_st, sizeof(KState) * (2 * nsys + 2), hipHostMallocDefault) !=
      hipSuccess || hipStreamSynchronize(nullptr) != hipSuccess)
    return TFS_ERR_NOMEM;
  {
    // the batch's SpMV only forms products (its reductions are per tile), so
    // the tuned shape cannot change any result
    StreamScope scope(A->stream);
    if ((rc = tune_spmv(A)) < 0) return rc;
  }
  return 0;
}

}  // namespace

namespace {

// The colour-ordered copy of a batch (see build_inner for one matrix): the
// block-diagonal S graph coloured greedily in S order, each system's rows
// sorted by colour, the systems kept in their order, so that the copy is a
// batch of the same systems whose sweeps have one level per colour.
int build_batch_inner(tfs_batch *B) {
  if (B->inner) return 0;
  tfs_matrix *A = B->A;
  const int n = A->n, ns = B->nsys;
  const long long nnz = A->nnz;
  std::vector<int> row(n + 1), col((size_t)nnz), src((size_t)nnz), orig(n);
  TFS_HIP(hipMemcpyAsync(row.data(), A->s_row, sizeof(int) * (n + 1), hipMemcpyDeviceToHost,
                         cur_stream()));
  TFS_HIP(hipMemcpyAsync(col.data(), A->s_col, sizeof(int) * nnz, hipMemcpyDeviceToHost,
                         cur_stream()));
  TFS_HIP(hipMemcpyAsync(src.data(), A->s_src, sizeof(int) * nnz, hipMemcpyDeviceToHost,
                         cur_stream()));
  TFS_HIP(hipMemcpyAsync(orig.data(), A->s_orig, sizeof(int) * n, hipMemcpyDeviceToHost,
                         cur_stream()));
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  std::vector<int> order(n);
  if (B->inner_order == TFS_INNER_BRICKS) {
    // Bricks per system, from the system's own caller numbering and size
    // (so that a system's copy is the same alone or in any batch), the
    // systems in their order: (system, colour, brick, caller order).
    std::vector<int> s_of(n);
    for (int k = 0; k < n; ++k) s_of[orig[k]] = k;
    int at = 0;
    for (int sy = 0; sy < ns; ++sy) {
      const int r0 = B->row_off[sy], m = B->n_of[sy];
      std::vector<int> crow(m + 1, 0), ccol, cdia(m, 0), pc, bstart, bcolour;
      for (int i = 0; i < m; ++i) {
        const int k = s_of[r0 + i];
        for (int e = row[k]; e < row[k + 1]; ++e) {  // the caller's entry order
          const int cj = orig[col[e]] - r0;
          if (cj == i) cdia[i] = (int)ccol.size();
          ccol.push_back(cj);
        }
        crow[i + 1] = (int)ccol.size();
      }
      if (m > 0) {
        const int edge = brick_edge(m);
        brick_colour_order(m, crow.data(), ccol.data(), cdia.data(), edge, edge * edge * edge, pc,
                           bstart, bcolour);
      }
      for (int i = 0; i < m; ++i) order[at++] = s_of[r0 + pc[i]];
    }
  } else {
    std::vector<int> perm, colour;
    colour_order(n, row.data(), col.data(), perm, &colour, B->inner_block);
    // system of each S row (from its caller row), then (system, colour, S) order
    std::vector<int> sys_of(n);
    for (int k = 0; k < n; ++k) {
      const int c = orig[k];
      sys_of[k] = (int)(std::upper_bound(B->row_off.begin(), B->row_off.end(), c) -
                        B->row_off.begin()) - 1;
    }
    for (int k = 0; k < n; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return sys_of[a] != sys_of[b] ? sys_of[a] < sys_of[b] : colour[a] < colour[b];
    });
  }
  std::vector<int> pos(n);
  for (int i = 0; i < n; ++i) pos[order[i]] = i;
  std::vector<int> prow(n + 1, 0), pcol((size_t)nnz), pdia(n, -1), emap((size_t)nnz);
  std::vector<std::pair<int, int>> ent;
  long long q = 0;
  for (int i = 0; i < n; ++i) {
    const int k0 = order[i];
    ent.clear();
    for (int k = row[k0]; k < row[k0 + 1]; ++k) ent.emplace_back(pos[col[k]], src[k]);
    std::sort(ent.begin(), ent.end());
    for (const auto &e : ent) {
      if (e.first == i) pdia[i] = (int)q;
      pcol[(size_t)q] = e.first;
      emap[(size_t)q] = e.second;
      ++q;
    }
    prow[i + 1] = (int)q;
  }
  auto *C = new tfs_batch;
  C->nsys = ns;
  C->n_of = B->n_of;
  C->row_off = B->row_off;
  // a system's entries stay one contiguous range in the copy
  long long eo = 0;
  for (int s = 0; s < ns; ++s) {
    C->nnz_of.push_back(B->nnz_of[s]);
    C->nnz_off.push_back(eo);
    eo += B->nnz_of[s];
  }
  Layout Ly;
  int rc = batch_build(C, n, nnz, prow.data(), pcol.data(), pdia.data(), Ly);
  if (rc == 0) rc = build_ell(C->A, Ly);  // the reliable-update loop's FP32 operator
  if (rc < 0) {
    tfs_batch_destroy(C);
    return rc;
  }
  std::vector<int> from_s(n);
  for (int i = 0; i < n; ++i) from_s[i] = order[Ly.s_orig[i]];
  if ((rc = upload(A, &B->inner_emap, emap)) < 0 || (rc = upload(A, &B->inner_from_s, from_s)) < 0) {
    tfs_batch_destroy(C);
    return rc;
  }
  C->A->persistent = A->persistent;
  C->A->dynamic = A->dynamic;
  C->A->coop = A->coop;
  B->inner = C;
  A->inner_ready = false;
  return 0;
}

}  // namespace

extern "C" int tfs_batch_create(int nsys, const int *n, const long long *nnz,
                                const int *const *row, const int *const *col,
                                const int *const *dia, int base, tfs_batch **out) {
  if (!out || nsys < 1 || !n || !nnz || !row || !col || !dia || (base != 0 && base != 1))
    return TFS_ERR_INVALID;
  *out = nullptr;
  // the combined block-diagonal structure (0-based)
  long long N = 0, NNZ = 0;
  for (int s = 0; s < nsys; ++s) {
    if (n[s] < 1 || nnz[s] < n[s]) return TFS_ERR_INVALID;
    N += n[s];
    NNZ += nnz[s];
  }
  if (N > 0x7ffffffeLL || NNZ > 0x7ffffffeLL) return TFS_ERR_INVALID;
  std::vector<int> cr(N + 1), cc((size_t)NNZ), cd((size_t)N);
  auto B = new tfs_batch;
  B->nsys = nsys;
  long long ro = 0, eo = 0;
  for (int s = 0; s < nsys; ++s) {
    B->n_of.push_back(n[s]);
    B->row_off.push_back((int)ro);
    B->nnz_of.push_back(nnz[s]);
    B->nnz_off.push_back(eo);
    if (row[s][0] - base != 0 || row[s][n[s]] - base != nnz[s]) {
      delete B;
      return TFS_ERR_INVALID;
    }
    for (int i = 0; i < n[s]; ++i) {
      cr[ro + i] = (int)(eo + row[s][i] - base);
      cd[ro + i] = (int)(eo + dia[s][i] - base);
    }
    for (long long k = 0; k < nnz[s]; ++k) {
      const int c = col[s][k] - base;
      if (c < 0 || c >= n[s]) {
        delete B;
        return TFS_ERR_INVALID;
      }
      cc[(size_t)(eo + k)] = (int)(ro + c);
    }
    ro += n[s];
    eo += nnz[s];
  }
  cr[N] = (int)NNZ;
  Layout Ly;
  const int rc = batch_build(B, N, NNZ, cr.data(), cc.data(), cd.data(), Ly);
  if (rc < 0) {
    tfs_batch_destroy(B);
    return rc;
  }
  *out = B;
  return TFS_OK;
}

extern "C" int tfs_batch_destroy(tfs_batch *B) {
  if (!B) return TFS_OK;
  if (B->inner) tfs_batch_destroy(B->inner);
  if (B->A && B->A->stream) (void)hipStreamSynchronize(B->A->stream);
  for (auto &g : B->graph)
    if (g) (void)hipGraphExecDestroy(g);
  if (B->h_st) (void)hipHostFree(B->h_st);
  tfs_matrix_destroy(B->A);
  delete B;
  return TFS_OK;
}

extern "C" int tfs_batch_set_values(tfs_batch *B, int sys, const double *val) {
  if (!B) return TFS_ERR_INVALID;
  StreamScope scope(B->A->stream);
  if (!B || sys < 0 || sys >= B->nsys || !val) return TFS_ERR_INVALID;
  TFS_HIP(hipMemcpyAsync(B->A->val + B->nnz_off[sys], val, sizeof(double) * B->nnz_of[sys],
                         hipMemcpyHostToDevice, cur_stream()));
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  values_changed(B->A);
  return TFS_OK;
}

extern "C" int tfs_batch_solve_device(tfs_batch *B, int prec, double *const *d_x,
                                      double *const *d_r1, const int *cap, const double *tol,
                                      int *niter, double *ini_res, double *fin_res,
                                      int *status) {
  if (!B) return TFS_ERR_INVALID;
  StreamScope scope(B->A->stream);
  if (!B || !d_x || !d_r1 || !cap || !tol || !niter || !ini_res || !fin_res || !status)
    return TFS_ERR_INVALID;
  if (prec < 0 || prec > 2) prec = 0;
  tfs_matrix *A = B->A;
  const int n = A->n, ns = B->nsys;
  TFS_TRY(ensure_vec64(A, 7));
  TFS_TRY(prec_form<double>(A, prec));
  const Vecs<double> V = vecs_of(A->vec64, n);
  KState *st = B->d_st, *all = B->d_st + 2 * ns;
  batch_init_states(B, 0, cap, tol);
  for (int s = 0; s < ns; ++s) B->h_st[s].ini_res = ini_res[s];
  TFS_TRY(batch_write(B));
  TFS_TRY(batch_in(B, d_x, A->stage_x, V.x));
  TFS_TRY(batch_in(B, d_r1, A->stage_r, V.r));
  const Tiles tl = tiles_of(B);
  // ||b|| and its early exit (per system); r = b - A x where not exited;
  // ||r|| and its early exit
  hipLaunchKernelGGL(k_sumsq_seg<double>, dim3(B->ntiles), dim3(kRedThreads), 0, cur_stream(), tl,
                     V.r, (int)EPI_BNRM, st, all, B->partials);
  finish_seg(B, EPI_BNRM, st, all);
  gather(n, V.r, (const int *)nullptr, V.q);
  launch_spmv<double, SPMV_RESID>(A, V.x, V.q, nullptr, nullptr, nullptr, all);
  hipLaunchKernelGGL(k_select_seg<double>, dim3(B->ntiles), dim3(kRedThreads), 0, cur_stream(),
                     tl, V.q, V.r, st, all);  // r = q for systems still running
  hipLaunchKernelGGL(k_sumsq_seg<double>, dim3(B->ntiles), dim3(kRedThreads), 0, cur_stream(), tl,
                     V.r, (int)EPI_INIRES, st, all, B->partials);
  finish_seg(B, EPI_INIRES, st, all);
  TFS_HIP(hipGetLastError());
  TFS_TRY(batch_read(B));
  int maxcap = 1;
  for (int s = 0; s < ns; ++s) {
    maxcap = std::max(maxcap, cap[s]);
    KState &h = B->h_st[s];
    if (!h.done && cap[s] < 1) {  // zero-trip loop (Cg.f90:97,181)
      h.done = 1;
      h.niter_out = 1;
      h.fin_res = h.err;
      h.status = kStNotConverged;
      B->h_st[2 * ns].ctr[0]++;
    }
  }
  if (B->h_st[2 * ns].ctr[0] == (unsigned)ns) B->h_st[2 * ns].done = 1;
  TFS_TRY(batch_write(B));
  TFS_TRY(batch_iterate<double>(B, prec, V, st, all, maxcap));
  TFS_TRY(batch_out(B, V.x, A->stage_x, d_x));
  TFS_TRY(batch_out(B, V.r, A->stage_r, d_r1));
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  TFS_TRY(check_sweep(A));
  int worst = TFS_OK;
  for (int s = 0; s < ns; ++s) {
    const KState &h = B->h_st[s];
    niter[s] = h.niter_out;
    fin_res[s] = h.fin_res;
    ini_res[s] = h.ini_res;
    status[s] = status_of(&h);
    if (status[s] == TFS_ERR_BREAKDOWN || (status[s] == TFS_NOT_CONVERGED && worst == TFS_OK))
      worst = status[s];
  }
  return worst;
}

// FP32-inner PCG in FP64 iterative refinement for every system at once
// (see solve_mixed_device); steps / inner iterations are per system.
namespace {

// One reliable-update iteration of every system on the batch's inner copy I
// (reliable_iteration per system): z = M^-1 r, rho, p, q = A32 p in
// difference form, (p, q), c += alpha p, r -= alpha q, ||r||.
void batch_reliable_iteration(tfs_batch *I, int prec, const Vecs<float> &F, double *c,
                              KState *st, KState *all) {
  tfs_matrix *A = I->A;
  const Tiles tl = tiles_of(I);
  const int nt = I->ntiles;
  if (prec == 2) dic_apply<float>(A, F.r, F.yL, F.zU, all);
  const int mode = prec == 2 ? 2 : prec == 1 ? 1 : 0;
  float *z = prec == 0 ? F.r : F.z;
  hipLaunchKernelGGL(k_rho_seg<float>, dim3(nt), dim3(kRedThreads), 0, cur_stream(), tl, F.r, F.z,
                     vals<float>(A).s_diag, F.zU, A->s_to_u, mode, st, all, I->partials);
  finish_seg(I, EPI_RHO, st, all);
  hipLaunchKernelGGL(k_pupdate_seg<float>, dim3(nt), dim3(kRedThreads), 0, cur_stream(), tl, z,
                     F.p, st, all);
  launch_ell<false>(A, F.p, F.q, all);
  hipLaunchKernelGGL(k_dot_seg<float>, dim3(nt), dim3(kRedThreads), 0, cur_stream(), tl, F.p, F.q,
                     st, all, I->partials);
  finish_seg(I, EPI_ALPHA, st, all);
  hipLaunchKernelGGL(k_axpy_mixed_seg, dim3(nt), dim3(kRedThreads), 0, cur_stream(), tl, c, F.r,
                     F.p, F.q, st, all, I->partials);
  finish_seg(I, EPI_NORM, st, all);
}

// The batch's reliable-update refinement (TFS_REFINE_RELIABLE, see
// solve_reliable_device): every system runs one FP32 CG recurrence on the
// colour-ordered inner copy, its directions summed into an FP64 correction;
// in rounds, each unconverged system runs until its recurrence residual has
// fallen by delta (its tiles idle once it has), then every such system's
// correction goes into x and its true FP64 residual replaces the recurrence's.
// Systems are independent: a system's path does not depend on the others.
int batch_solve_reliable(tfs_batch *B, int prec, double *const *d_x, double *const *d_r1, int cap,
                         const double *tol, double delta, int *steps, int *inner_iters,
                         double *ini_res, double *fin_res, int *status) {
  if (!(delta > 0.0)) delta = 0.1;
  delta = std::min(0.5, std::max(1e-6, delta));
  tfs_matrix *A = B->A;
  const int n = A->n, ns = B->nsys;
  TFS_TRY(build_batch_inner(B));
  tfs_batch *I = B->inner;
  if (!A->inner_ready) {
    gather((int)A->nnz, A->val, B->inner_emap, I->A->val);
    values_changed(I->A);
    A->inner_ready = true;
  }
  const int *to_s = B->inner_from_s;  // I's S position -> B's
  TFS_TRY(ensure_vec64(A, 8));
  TFS_TRY(ensure_vec32(I->A, 7));
  TFS_TRY(prepare_values<double>(A));
  TFS_TRY(prec_form<float>(I->A, prec));
  TFS_TRY(prepare_ell(I->A));
  const Vecs<double> D = vecs_of(A->vec64, n);
  double *bS = D.r2, *c = D.z;  // c: FP64 correction in I's S order
  const Vecs<float> F = vecs_of(I->A->vec32, n);
  KState *so = B->d_st + ns, *allo = B->d_st + 2 * ns + 1;
  KState *si = I->d_st, *alli = I->d_st + 2 * ns;
  KState *hi = I->h_st, *ho = B->h_st + ns;
  const Tiles tl = tiles_of(B);
  const int g = vgrid(n);
  TFS_TRY(batch_in(B, d_x, A->stage_x, D.x));
  TFS_TRY(batch_in(B, d_r1, A->stage_r, bS));  // b
  std::vector<int> caps(ns, cap), stalls(ns, 0), updates(ns, 0), st_out(ns, TFS_OK);
  std::vector<double> err(ns), err0(ns), best(ns);
  auto true_residual = [&]() -> int {
    batch_init_states(B, ns, caps.data(), tol);
    TFS_TRY(batch_write(B));
    gather(n, bS, (const int *)nullptr, D.r);
    launch_spmv<double, SPMV_RESID>(A, D.x, D.r, nullptr, nullptr, nullptr, nullptr);
    hipLaunchKernelGGL(k_sumsq_seg<double>, dim3(B->ntiles), dim3(kRedThreads), 0, cur_stream(),
                       tl, D.r, (int)EPI_BNRM, so, allo, B->partials);
    finish_seg(B, EPI_BNRM, so, allo);
    TFS_HIP(hipGetLastError());
    return batch_read(B);
  };
  TFS_TRY(true_residual());
  for (int s = 0; s < ns; ++s) err[s] = err0[s] = best[s] = ho[s].bnrm2;
  hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, c, 0.0);
  std::vector<char> active(ns, 0), started(ns, 0);
  for (;;) {
    int nact = 0;
    for (int s = 0; s < ns; ++s) {
      active[s] = err[s] >= tol[s] && st_out[s] == TFS_OK;
      if (active[s] && updates[s] >= cap) {
        st_out[s] = TFS_NOT_CONVERGED;
        active[s] = 0;
      }
      nact += active[s];
    }
    if (!nact) break;
    // each active system's recurrence continues from its true residual
    gather(n, D.r, to_s, F.r);
    int ndone = 0;
    for (int s = 0; s < ns; ++s) {
      KState &h = hi[s];
      if (!started[s]) {
        std::memset(&h, 0, sizeof(KState));
        h.cap = kSegmentCap;
      } else {
        h.rho_old = h.rho;  // the last iteration's rho (kept on its exit)
        h.cap = h.iter + kSegmentCap;
      }
      h.done = active[s] ? 0 : 1;
      h.status = 0;
      h.tol = std::max(delta * err[s], 0.5 * tol[s]);
      h.err = err[s];
      if (active[s]) started[s] = 1;
      ndone += !active[s];
    }
    KState &ha = hi[2 * ns];
    std::memset(&ha, 0, sizeof(KState));
    ha.cap = ns;
    ha.ctr[0] = (unsigned)ndone;
    TFS_TRY(batch_write(I));
    TFS_TRY(batch_iterate_fn(I, 2, prec, si, alli, 1 << 30,
                             [&] { batch_reliable_iteration(I, prec, F, c, si, alli); }));
    // x += c for every system (c is zero for the idle ones); true residuals
    hipLaunchKernelGGL(k_fold_correction, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, D.x, c,
                       to_s);
    TFS_TRY(true_residual());
    for (int s = 0; s < ns; ++s) {
      if (!active[s]) continue;
      ++updates[s];
      const double e = ho[s].bnrm2;
      err[s] = e;
      if (hi[s].status == kStBreakdown || !std::isfinite(e)) {
        st_out[s] = TFS_ERR_BREAKDOWN;
      } else if (hi[s].status == kStNotConverged && e >= tol[s]) {
        st_out[s] = TFS_NOT_CONVERGED;  // a segment ran out of iterations
      } else {
        stalls[s] = e >= best[s] ? stalls[s] + 1 : 0;
        best[s] = std::min(best[s], e);
        if (stalls[s] >= 3 && e >= tol[s]) st_out[s] = TFS_NOT_CONVERGED;
      }
    }
  }
  TFS_TRY(batch_out(B, D.x, A->stage_x, d_x));
  TFS_TRY(batch_out(B, D.r, A->stage_r, d_r1));
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  TFS_TRY(check_sweep(A));
  TFS_TRY(check_sweep(I->A));
  int worst = TFS_OK;
  for (int s = 0; s < ns; ++s) {
    ini_res[s] = err0[s];
    fin_res[s] = err[s];
    status[s] = st_out[s];
    steps[s] = st_out[s] == TFS_NOT_CONVERGED ? cap + 1 : updates[s];
    inner_iters[s] = started[s] ? hi[s].iter : 0;
    if (status[s] == TFS_ERR_BREAKDOWN || (status[s] == TFS_NOT_CONVERGED && worst == TFS_OK))
      worst = status[s];
  }
  return worst;
}

}  // namespace

extern "C" int tfs_batch_solve_mixed_device(tfs_batch *B, int prec, double *const *d_x,
                                            double *const *d_r1, int cap, const double *tol,
                                            double inner_rtol, int *steps, int *inner_iters,
                                            double *ini_res, double *fin_res, int *status) {
  if (!B) return TFS_ERR_INVALID;
  StreamScope scope(B->A->stream);
  if (!B || !d_x || !d_r1 || !tol || !steps || !inner_iters || !ini_res || !fin_res || !status)
    return TFS_ERR_INVALID;
  if (prec < 0 || prec > 2) prec = 0;
  const bool copy_order =
      B->inner_order == TFS_INNER_COLOURED || B->inner_order == TFS_INNER_BRICKS;
  if (B->refine == TFS_REFINE_RELIABLE && copy_order)
    return batch_solve_reliable(B, prec, d_x, d_r1, cap, tol, inner_rtol, steps, inner_iters,
                                ini_res, fin_res, status);
  if (!(inner_rtol > 0.0)) inner_rtol = 1e-2;
  inner_rtol = std::min(0.5, std::max(1e-6, inner_rtol));
  tfs_matrix *A = B->A;
  const int n = A->n, ns = B->nsys;
  // the inner solves run on I: this batch, or its colour-ordered copy (prec 2)
  const bool coloured = copy_order && prec == 2;
  tfs_batch *I = B;
  if (coloured) {
    TFS_TRY(build_batch_inner(B));
    I = B->inner;
    if (!A->inner_ready) {
      gather((int)A->nnz, A->val, B->inner_emap, I->A->val);
      values_changed(I->A);
      A->inner_ready = true;
    }
  }
  const int *to_s = coloured ? B->inner_from_s : nullptr;  // I's S position -> B's
  TFS_TRY(ensure_vec64(A, 8));
  TFS_TRY(ensure_vec32(I->A, 7));
  TFS_TRY(prepare_values<double>(A));
  TFS_TRY(prec_form<float>(I->A, prec));
  // FP64: x, r, z (inner result), p, q = A p, r2 slot = b (the sweep slots
  // stay armed for FP64 solves)
  const Vecs<double> D = vecs_of(A->vec64, n);
  double *bS = D.r2;
  const bool krylov = B->refine == TFS_REFINE_FCG;
  const Vecs<float> F = vecs_of(I->A->vec32, n);
  KState *so = B->d_st + ns, *allo = B->d_st + 2 * ns + 1;
  KState *si = I->d_st, *alli = I->d_st + 2 * ns;
  KState *hi = I->h_st, *ho = B->h_st + ns;
  const Tiles tl = tiles_of(B), tli = tiles_of(I);
  const int g = vgrid(n);
  TFS_TRY(batch_in(B, d_x, A->stage_x, D.x));
  TFS_TRY(batch_in(B, d_r1, A->stage_r, bS));  // b
  std::vector<int> caps(ns, cap), stalls(ns, 0), outer(ns, 0), st_out(ns, TFS_OK);
  std::vector<double> err(ns), err0(ns);
  for (int s = 0; s < ns; ++s) inner_iters[s] = 0;
  auto outer_residual = [&]() -> int {
    batch_init_states(B, ns, caps.data(), tol);
    TFS_TRY(batch_write(B));
    gather(n, bS, (const int *)nullptr, D.r);
    launch_spmv<double, SPMV_RESID>(A, D.x, D.r, nullptr, nullptr, nullptr, nullptr);
    hipLaunchKernelGGL(k_sumsq_seg<double>, dim3(B->ntiles), dim3(kRedThreads), 0, cur_stream(),
                       tl, D.r, (int)EPI_BNRM, so, allo, B->partials);
    finish_seg(B, EPI_BNRM, so, allo);
    TFS_HIP(hipGetLastError());
    return batch_read(B);
  };
  TFS_TRY(outer_residual());
  for (int s = 0; s < ns; ++s) err[s] = err0[s] = ho[s].bnrm2;
  std::vector<char> active(ns, 0), have_p(ns, 0);
  std::vector<double> pq_prev(ns, 0.0);
  // per-system scalars into the outer states (read, edit, write back)
  auto dots = [&](auto &&launch) -> int {
    launch();
    TFS_HIP(hipGetLastError());
    return batch_read(B);
  };
  for (;;) {
    int nact = 0;
    for (int s = 0; s < ns; ++s) {
      active[s] = err[s] >= tol[s] && st_out[s] == TFS_OK;
      if (active[s] && outer[s] >= cap) {
        st_out[s] = TFS_NOT_CONVERGED;
        active[s] = 0;
      }
      nact += active[s];
    }
    if (!nact) break;
    // inner FP32 solves of the active systems; the others are done at once
    std::vector<int> icap(ns, 100000);
    std::vector<double> itol(ns);
    for (int s = 0; s < ns; ++s)
      itol[s] = krylov ? inner_rtol * err[s] : std::max(inner_rtol * err[s], 0.5 * tol[s]);
    batch_init_states(I, 0, icap.data(), itol.data());
    int ndone = 0;
    for (int s = 0; s < ns; ++s)
      if (!active[s]) {
        hi[s].done = 1;
        ++ndone;
      }
    hi[2 * ns].ctr[0] = (unsigned)ndone;
    TFS_TRY(batch_write(I));
    gather(n, D.r, to_s, F.r);
    hipLaunchKernelGGL(k_fill<float>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, F.x, 0.0f);
    hipLaunchKernelGGL(k_sumsq_seg<float>, dim3(I->ntiles), dim3(kRedThreads), 0, cur_stream(),
                       tli, F.r, (int)EPI_INIRES, si, alli, I->partials);
    finish_seg(I, EPI_INIRES, si, alli);
    TFS_HIP(hipGetLastError());
    TFS_TRY(batch_read(I));
    TFS_TRY(batch_iterate<float>(I, prec, F, si, alli, 100000));
    for (int s = 0; s < ns; ++s)
      if (active[s]) {
        ++outer[s];
        inner_iters[s] += std::min(hi[s].niter_out, 100000);
        if (hi[s].status == kStBreakdown) st_out[s] = TFS_ERR_BREAKDOWN;
      }
    std::vector<double> e_new(ns, 0.0);
    if (!krylov) {
      // defect correction: x += z; r = b - A x
      hipLaunchKernelGGL(k_add_correction, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, D.x,
                         F.x, to_s);
      TFS_TRY(outer_residual());
      for (int s = 0; s < ns; ++s) e_new[s] = ho[s].bnrm2;
    } else {
      // flexible CG per system (see solve_mixed_device): the outer states of
      // the inactive systems are marked done, so their tiles are skipped
      for (int s = 0; s < ns; ++s) {
        ho[s].done = active[s] ? 0 : 1;
        ho[s].iter = have_p[s] ? 1 : 0;   // k_pupdate_seg: p = z on a system's first step
      }
      TFS_TRY(batch_write(B));
      hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, D.z, 0.0);
      hipLaunchKernelGGL(k_add_correction, dim3(g), dim3(kRedThreads), 0, cur_stream(), n, D.z,
                         F.x, to_s);
      TFS_TRY(dots([&] {
        hipLaunchKernelGGL(k_dot_seg<double>, dim3(B->ntiles), dim3(kRedThreads), 0,
                           cur_stream(), tl, D.z, D.q, so, allo, B->partials);
        finish_seg(B, EPI_AUX0, so, allo);  // (z, A p_prev)
      }));
      for (int s = 0; s < ns; ++s)
        if (active[s] && have_p[s]) ho[s].beta = -ho[s].aux[0] / pq_prev[s];
      TFS_TRY(batch_write(B));
      hipLaunchKernelGGL(k_pupdate_seg<double>, dim3(B->ntiles), dim3(kRedThreads), 0,
                         cur_stream(), tl, D.z, D.p, so, allo);
      launch_spmv<double, SPMV_PLAIN>(A, D.p, D.q, nullptr, nullptr, nullptr, allo);
      TFS_TRY(dots([&] {
        hipLaunchKernelGGL(k_dot_seg<double>, dim3(B->ntiles), dim3(kRedThreads), 0,
                           cur_stream(), tl, D.p, D.q, so, allo, B->partials);
        finish_seg(B, EPI_AUX1, so, allo);  // (p, A p)
        hipLaunchKernelGGL(k_dot_seg<double>, dim3(B->ntiles), dim3(kRedThreads), 0,
                           cur_stream(), tl, D.p, D.r, so, allo, B->partials);
        finish_seg(B, EPI_AUX2, so, allo);  // (p, r)
      }));
      for (int s = 0; s < ns; ++s) {
        if (!active[s]) continue;
        const double pq = ho[s].aux[1], alpha = ho[s].aux[2] / pq;
        if (!(pq > 0.0) || !std::isfinite(alpha)) {
          st_out[s] = TFS_ERR_BREAKDOWN;
          ho[s].done = 1;
          continue;
        }
        ho[s].alpha = alpha;
        pq_prev[s] = pq;
        have_p[s] = 1;
      }
      TFS_TRY(batch_write(B));
      TFS_TRY(dots([&] {
        hipLaunchKernelGGL(k_axpy_norm_seg<double>, dim3(B->ntiles), dim3(kRedThreads), 0,
                           cur_stream(), tl, D.x, D.r, D.p, D.q, so, allo, B->partials);
        finish_seg(B, EPI_NAUX3, so, allo);  // ||r|| of the recurrence
      }));
      bool confirm = false;
      for (int s = 0; s < ns; ++s) {
        e_new[s] = active[s] ? ho[s].aux[3] : err[s];
        confirm |= active[s] && e_new[s] < tol[s];
      }
      if (confirm) {
        // the true residual of each system whose recurrence passed its tol,
        // formed in the z slot: it confirms convergence, or the system
        // restarts from it.  Only those systems are touched, so no system's
        // path depends on the others.
        std::vector<char> check(ns, 0), restart(ns, 0);
        for (int s = 0; s < ns; ++s) {
          check[s] = active[s] && e_new[s] < tol[s];
          ho[s].done = check[s] ? 0 : 1;
        }
        TFS_TRY(batch_write(B));
        gather(n, bS, (const int *)nullptr, D.z);
        launch_spmv<double, SPMV_RESID>(A, D.x, D.z, nullptr, nullptr, nullptr, nullptr);
        TFS_TRY(dots([&] {
          hipLaunchKernelGGL(k_sumsq_seg<double>, dim3(B->ntiles), dim3(kRedThreads), 0,
                             cur_stream(), tl, D.z, (int)EPI_NAUX4, so, allo, B->partials);
          finish_seg(B, EPI_NAUX4, so, allo);
        }));
        bool any = false;
        for (int s = 0; s < ns; ++s) {
          if (!check[s]) continue;
          e_new[s] = ho[s].aux[4];
          restart[s] = e_new[s] >= tol[s];
          any |= restart[s];
        }
        if (any) {
          for (int s = 0; s < ns; ++s) {
            ho[s].done = restart[s] ? 0 : 1;
            if (restart[s]) have_p[s] = 0;
          }
          TFS_TRY(batch_write(B));
          hipLaunchKernelGGL(k_select_seg<double>, dim3(B->ntiles), dim3(kRedThreads), 0,
                             cur_stream(), tl, D.z, D.r, so, allo);
          TFS_HIP(hipGetLastError());
        }
      }
    }
    for (int s = 0; s < ns; ++s) {
      if (!active[s]) continue;
      const double e = e_new[s];
      if (!std::isfinite(e)) st_out[s] = TFS_ERR_BREAKDOWN;
      stalls[s] = e > (krylov ? 1.0 : 0.9) * err[s] ? stalls[s] + 1 : 0;
      if (stalls[s] >= 3 && e >= tol[s]) st_out[s] = TFS_NOT_CONVERGED;
      err[s] = e;
    }
  }
  if (krylov) {
    // the returned residuals are the true ones
    TFS_TRY(outer_residual());
    for (int s = 0; s < ns; ++s)
      if (st_out[s] == TFS_OK) err[s] = ho[s].bnrm2;
  }
  TFS_TRY(batch_out(B, D.x, A->stage_x, d_x));
  TFS_TRY(batch_out(B, D.r, A->stage_r, d_r1));
  TFS_HIP(hipGetLastError());
  TFS_HIP(hipStreamSynchronize(cur_stream()));
  TFS_TRY(check_sweep(A));
  if (I != B) TFS_TRY(check_sweep(I->A));
  int worst = TFS_OK;
  for (int s = 0; s < ns; ++s) {
    ini_res[s] = err0[s];
    fin_res[s] = err[s];
    status[s] = st_out[s];
    steps[s] = st_out[s] == TFS_NOT_CONVERGED ? cap + 1 : outer[s];
    if (status[s] == TFS_ERR_BREAKDOWN || (status[s] == TFS_NOT_CONVERGED && worst == TFS_OK))
      worst = status[s];
  }
  return worst;
}

extern "C" int tfs_batch_set_inner_block(tfs_batch *B, int rows) {
  if (!B || rows < 1 || rows > 4096) return TFS_ERR_INVALID;
  if (rows == B->inner_block) return TFS_OK;
  B->inner_block = rows;
  if (B->inner) {  // rebuilt on the next refinement solve
    StreamScope scope(B->A->stream);
    tfs_batch_destroy(B->inner);
    B->inner = nullptr;
    TFS_TRY(dev_free(B->A, B->inner_emap, sizeof(int) * (size_t)B->A->nnz));
    TFS_TRY(dev_free(B->A, B->inner_from_s, sizeof(int) * (size_t)B->A->n));
    B->inner_emap = B->inner_from_s = nullptr;
  }
  return TFS_OK;
}

extern "C" int tfs_batch_set_refinement(tfs_batch *B, int outer) {
  if (!B || (outer != TFS_REFINE_FCG && outer != TFS_REFINE_DEFECT && outer != TFS_REFINE_RELIABLE))
    return TFS_ERR_INVALID;
  B->refine = outer;
  return TFS_OK;
}

extern "C" int tfs_batch_set_inner_order(tfs_batch *B, int order) {
  if (!B || (order != TFS_INNER_REFERENCE && order != TFS_INNER_COLOURED &&
             order != TFS_INNER_BRICKS))
    return TFS_ERR_INVALID;
  if (order == B->inner_order) return TFS_OK;
  B->inner_order = order;
  if (B->inner) {  // rebuilt on the next refinement solve
    StreamScope scope(B->A->stream);
    tfs_batch_destroy(B->inner);
    B->inner = nullptr;
    TFS_TRY(dev_free(B->A, B->inner_emap, sizeof(int) * (size_t)B->A->nnz));
    TFS_TRY(dev_free(B->A, B->inner_from_s, sizeof(int) * (size_t)B->A->n));
    B->inner_emap = B->inner_from_s = nullptr;
  }
  return TFS_OK;
}

extern "C" int tfs_batch_get_info(const tfs_batch *B, int *nsys, int *n_total,
                                  long long *nnz_total, int *graph_launches,
                                  int *spmv_rows, int *spmv_waves) {
  if (!B) return TFS_ERR_INVALID;
  if (nsys) *nsys = B->nsys;
  if (n_total) *n_total = B->A->n;
  if (nnz_total) *nnz_total = B->A->nnz;
  if (graph_launches) *graph_launches = B->graph_launches;
  if (spmv_rows) *spmv_rows = B->A->spmv_rows;
  if (spmv_waves) *spmv_waves = B->A->spmv_waves;
  return TFS_OK;
}

extern "C" int tfs_matrix_set_sweep_schedule(tfs_matrix *A, int schedule) {
  if (!A || schedule < TFS_SWEEP_LEVELS || schedule > TFS_SWEEP_COOPERATIVE)
    return TFS_ERR_INVALID;
  StreamScope scope(A->stream);
  TFS_HIP(hipStreamSynchronize(A->stream));
  const bool was_levels = !A->persistent;
  A->persistent = schedule != TFS_SWEEP_LEVELS;
  A->dynamic = schedule == TFS_SWEEP_DYNAMIC;
  A->coop = schedule == TFS_SWEEP_COOPERATIVE;
  A->sweep_fixed = true;
  if (was_levels && A->persistent) {
    // the persistent sweeps need their outputs armed
    if (A->vec64) arm(A->n, A->vec64 + 2 * (size_t)A->n);
    if (A->vec32) arm(A->n, A->vec32 + 2 * (size_t)A->n);
    TFS_HIP(hipStreamSynchronize(A->stream));
  }
  if (A->inner) return tfs_matrix_set_sweep_schedule(A->inner, schedule);
  return TFS_OK;
}
