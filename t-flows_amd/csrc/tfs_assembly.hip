// tfs_assembly.hip -- device-side assembly of A % val from face coefficients.
//
// T-FlowS forms its pressure and momentum matrices in a loop over faces that
// scatters through the topology's pos/dia arrays
// (Sources/Process/Pressure_Matrix_Fractional.f90:19-48,
//  Sources/Process/NewUVW.f90:349-384 for the faces, :500-506 for the
//  inertia term, :586-590 for under-relaxation; pos/dia from
//  Sources/Process/Matrix_Mod/Topology.f90:105-121):
//
//   val := 0
//   do s = 1, n_faces
//     if (c2 > 0) then  val(pos(1,s)) -= a12;  val(dia(c1)) += a12
//                       val(pos(2,s)) -= a21;  val(dia(c2)) += a21
//     else (boundary kinds that contribute)    val(dia(c1)) += a12
//   do c: val(dia(c)) += cell(c);  sav(c) = val(dia(c));  val(dia(c)) /= urf
//
// Every off-diagonal entry belongs to one face, so it is 0 - a12 whatever the
// order.  A diagonal receives one addition per incident face, in face order;
// here each cell sums its own faces -- kept sorted by face index -- so the
// result is bitwise that of the sequential loop, with one thread per cell and
// no atomics.  The values then go through the solver's usual change path
// (layouts regathered, factor rebuilt on the next solve).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "tfs_internal.hpp"
#include "tfs_solver.h"

namespace {

constexpr int kThreads = 256;

#define TFS_A_HIP(expr)                        \
  do {                                         \
    if ((expr) != hipSuccess) {                \
      (void)hipGetLastError();                 \
      return TFS_ERR_DEVICE;                   \
    }                                          \
  } while (0)

// Off-diagonals: one thread per interior face.  val has been zeroed, and the
// reference's val(pos) = val(pos) - a is 0 - a (which keeps the sign of a
// zero coefficient as the reference does).
__global__ void __launch_bounds__(kThreads)
k_faces(long long nf, const int *__restrict__ c2, const int *__restrict__ pos,
        const double *__restrict__ a12, const double *__restrict__ a21, double *__restrict__ val) {
  for (long long s = blockIdx.x * (long long)kThreads + threadIdx.x; s < nf;
       s += (long long)gridDim.x * kThreads) {
    if (c2[s] < 0) continue;
    const double u = a12[s], w = a21 ? a21[s] : u;
    val[pos[2 * s]] = 0.0 - u;
    val[pos[2 * s + 1]] = 0.0 - w;
  }
}

// Diagonals: one thread per cell, its faces in ascending face order.  Entry
// 2 s + k of the incidence list means "face s, this cell is c(k+1)"; the
// byte beside it is set for an interior face, so the face's own cells are
// not gathered (the list and its bytes are contiguous per cell; the
// coefficients are the only gathers).
__global__ void __launch_bounds__(kThreads)
k_cells(int n, const int *__restrict__ ptr, const long long *__restrict__ inc,
        const unsigned char *__restrict__ interior, const double *__restrict__ a12,
        const double *__restrict__ a21, const unsigned char *__restrict__ bnd_on,
        const double *__restrict__ cell, double urf, const int *__restrict__ dia,
        double *__restrict__ val, double *__restrict__ sav) {
  for (int c = blockIdx.x * kThreads + threadIdx.x; c < n; c += gridDim.x * kThreads) {
    double d = 0.0;
    for (int k = ptr[c]; k < ptr[c + 1]; ++k) {
      const long long e = inc[k], s = e >> 1;
      if (e & 1) {
        d = d + (a21 ? a21[s] : a12[s]);
      } else if (interior[k] || !bnd_on || bnd_on[s]) {
        d = d + a12[s];
      }
    }
    if (cell) d = d + cell[c];
    if (sav) sav[c] = d;
    if (urf != 1.0) d = d / urf;
    val[dia[c]] = d;
  }
}

__global__ void __launch_bounds__(kThreads) k_zero(long long m, double *__restrict__ v) {
  for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < m;
       i += (long long)gridDim.x * kThreads)
    v[i] = 0.0;
}

int grid_of(long long m) { return (int)std::max(1LL, std::min(4096LL, (m + kThreads - 1) / kThreads)); }

}  // namespace

struct tfs_assembly {
  tfs_matrix *A = nullptr;
  int n = 0;
  long long nf = 0, nnz = 0;
  int *c2 = nullptr, *pos = nullptr, *dia = nullptr, *ptr = nullptr;
  long long *inc = nullptr;
  unsigned char *interior = nullptr;   // per incidence: the face is interior
};

extern "C" int tfs_assembly_destroy(tfs_assembly *S) {
  if (!S) return TFS_OK;
  for (void *p : {(void *)S->c2, (void *)S->pos, (void *)S->dia, (void *)S->ptr, (void *)S->inc,
                  (void *)S->interior})
    if (p) (void)hipFree(p);
  delete S;
  return TFS_OK;
}

extern "C" int tfs_assembly_create(tfs_matrix *A, long long n_faces, const int *faces_c,
                                   const int *pos, const int *dia, int base, tfs_assembly **out) {
  if (!out) return TFS_ERR_INVALID;
  *out = nullptr;
  tfs_matrix_info info;
  if (!A || tfs_matrix_get_info(A, &info) != TFS_OK || n_faces < 0 ||
      (n_faces > 0 && (!faces_c || !pos)) || (info.n > 0 && !dia) || (base != 0 && base != 1))
    return TFS_ERR_INVALID;
  const int n = info.n;
  const long long nf = n_faces, nnz = info.nnz;
  // 0-based copies; c2 < 0 marks a boundary face (the reference's c2 <= 0 with 1-based cells)
  std::vector<int> hc2((size_t)nf), hpos((size_t)(2 * nf)), hdia((size_t)n);
  std::vector<int> cnt((size_t)n + 1, 0);
  for (long long s = 0; s < nf; ++s) {
    const int c1 = faces_c[2 * s] - base, c2 = faces_c[2 * s + 1] - base;
    if (c1 < 0 || c1 >= n || c2 >= n || c2 == c1) return TFS_ERR_INVALID;
    hc2[(size_t)s] = c2 >= 0 ? c2 : -1;
    cnt[(size_t)c1 + 1]++;
    if (c2 >= 0) {
      cnt[(size_t)c2 + 1]++;
      for (int k = 0; k < 2; ++k) {
        const long long q = (long long)pos[2 * s + k] - base;
        if (q < 0 || q >= nnz) return TFS_ERR_INVALID;
        hpos[(size_t)(2 * s + k)] = (int)q;
      }
    } else {
      hpos[(size_t)(2 * s)] = hpos[(size_t)(2 * s + 1)] = 0;
    }
  }
  for (int c = 0; c < n; ++c) {
    hdia[(size_t)c] = dia[c] - base;
    if (hdia[(size_t)c] < 0 || hdia[(size_t)c] >= nnz) return TFS_ERR_INVALID;
  }
  // cell -> incident faces, ascending face order (a counting sort by cell
  // over the faces in order keeps each cell's list sorted)
  for (int c = 0; c < n; ++c) cnt[(size_t)c + 1] += cnt[(size_t)c];
  std::vector<long long> hinc((size_t)cnt[(size_t)n]);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (long long s = 0; s < nf; ++s) {
    const int c1 = faces_c[2 * s] - base, c2 = hc2[(size_t)s];
    hinc[(size_t)fill[(size_t)c1]++] = 2 * s;
    if (c2 >= 0) hinc[(size_t)fill[(size_t)c2]++] = 2 * s + 1;
  }
  std::vector<unsigned char> hint(hinc.size());
  for (size_t k = 0; k < hinc.size(); ++k) hint[k] = hc2[(size_t)(hinc[k] >> 1)] >= 0 ? 1 : 0;
  auto *S = new tfs_assembly;
  S->A = A;
  S->n = n;
  S->nf = nf;
  S->nnz = nnz;
  auto up = [](void **dst, const void *src, size_t bytes) -> bool {
    if (hipMalloc(dst, std::max<size_t>(bytes, 8)) != hipSuccess) return false;
    return bytes == 0 || hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  if (!up((void **)&S->c2, hc2.data(), sizeof(int) * hc2.size()) ||
      !up((void **)&S->pos, hpos.data(), sizeof(int) * hpos.size()) ||
      !up((void **)&S->dia, hdia.data(), sizeof(int) * hdia.size()) ||
      !up((void **)&S->ptr, cnt.data(), sizeof(int) * cnt.size()) ||
      !up((void **)&S->inc, hinc.data(), sizeof(long long) * hinc.size()) ||
      !up((void **)&S->interior, hint.data(), hint.size())) {
    (void)hipGetLastError();
    tfs_assembly_destroy(S);
    return TFS_ERR_NOMEM;
  }
  *out = S;
  return TFS_OK;
}

extern "C" int tfs_assemble_device(tfs_assembly *S, const double *d_a12, const double *d_a21,
                                   const unsigned char *d_bnd_on, const double *d_cell, double urf,
                                   double *d_sav) {
  if (!S || (S->nf > 0 && !d_a12) || !(urf > 0.0)) return TFS_ERR_INVALID;
  double *val = tfs_matrix_values_device(S->A);
  hipStream_t st = tfs::matrix_stream(S->A);
  hipLaunchKernelGGL(k_zero, dim3(grid_of(S->nnz)), dim3(kThreads), 0, st, S->nnz, val);
  if (S->nf > 0)
    hipLaunchKernelGGL(k_faces, dim3(grid_of(S->nf)), dim3(kThreads), 0, st, S->nf, S->c2, S->pos,
                       d_a12, d_a21, val);
  if (S->n > 0)
    hipLaunchKernelGGL(k_cells, dim3(grid_of(S->n)), dim3(kThreads), 0, st, S->n, S->ptr, S->inc,
                       S->interior, d_a12, d_a21, d_bnd_on, d_cell, urf, S->dia, val, d_sav);
  TFS_A_HIP(hipGetLastError());
  TFS_A_HIP(hipStreamSynchronize(st));
  return tfs_matrix_values_changed(S->A);
}
