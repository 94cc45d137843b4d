// tfs_topology.hip -- the CRS topology of a face list, built on the device.
//
// Sources/Process/Matrix_Mod/Topology.f90 forms the matrix structure from the
// faces in four sequential loops: stencil widths (:34-41), row starts with the
// diagonal first (:54-59), off-diagonal columns appended in face order
// (:62-71), each row sorted ascending with its diagonal located (:75-82), and
// the two entries of every interior face located (:106-121).  Here:
//
//   k_width   one thread per face, atomic width counts per cell; the count an
//             atomic returns is the face's place in that row, kept per face;
//   k_scan_*  row starts: per-block sums, one scan of the block sums, per-block
//             rescan (an exclusive scan of width + 1 with 64-bit totals);
//   k_fill    one thread per face, no atomics: each face's two entries go to
//             the places k_width returned, each carrying a tag naming its face
//             and side (2 s + side);
//   k_sort    one thread per cell: insertion sort of its row's (column, tag)
//             pairs by column, the diagonal's position (the last entry equal
//             to the cell, as the reference's loop leaves it), and pos: every
//             tag in a run of equal columns gets the run's last position (the
//             reference's search keeps the last match).
//
// Atomics make the fill order nondeterministic, but a row holds a multiset of
// columns and sorting it fixes the order, so the result is bitwise the
// sequential one; tags only travel with their columns.  Everything is
// integer traffic: the kernels are bound by the device-scope atomics (the
// width counts land on cells in mesh order, which is random on an
// unstructured numbering) and by the scattered column and pos writes, not by
// arithmetic.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "tfs_solver.h"

namespace {

constexpr int kThreads = 256;
constexpr int kScanItems = 4;                       // per thread
constexpr int kScanTile = kThreads * kScanItems;    // cells per scan block

#define TFS_T_HIP(expr)        \
  do {                         \
    if ((expr) != hipSuccess) { \
      (void)hipGetLastError(); \
      rc = TFS_ERR_DEVICE;     \
      goto done;               \
    }                          \
  } while (0)

int grid_of(long long m) {
  long long g = (m + kThreads - 1) / kThreads;
  return (int)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

__global__ void __launch_bounds__(kThreads) k_fill_int(long long m, int v, int *__restrict__ a) {
  for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < m;
       i += (long long)gridDim.x * kThreads)
    a[i] = v;
}

// widths of the off-diagonal part; err gets 1 for c1 outside 1..n or c2 > n
// (c2 <= 0 is a boundary face, as the reference's test c2 > 0 has it)
__global__ void __launch_bounds__(kThreads)
k_width(int n, long long nf, const int *__restrict__ faces, int *__restrict__ width,
        int2 *__restrict__ slot, int *__restrict__ err) {
  for (long long s = blockIdx.x * (long long)kThreads + threadIdx.x; s < nf;
       s += (long long)gridDim.x * kThreads) {
    const int2 f = reinterpret_cast<const int2 *>(faces)[s];
    if (f.x < 1 || f.x > n || f.y > n) {
      atomicOr(err, 1);
      continue;
    }
    if (f.y > 0) {
      // the count before this face is the face's place in each row, after
      // the diagonal
      const int k1 = atomicAdd(&width[f.x - 1], 1);
      const int k2 = atomicAdd(&width[f.y - 1], 1);
      slot[s] = make_int2(k1 + 1, k2 + 1);
    }
  }
}

// block sums of (width + 1) over tiles of kScanTile cells
__global__ void __launch_bounds__(kThreads)
k_scan_sums(int n, const int *__restrict__ width, long long *__restrict__ bsum) {
  __shared__ long long part[kThreads / 64];
  const long long base = (long long)blockIdx.x * kScanTile;
  long long t = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const long long c = base + (long long)k * kThreads + threadIdx.x;
    if (c < n) t += width[c] + 1;
  }
  for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long s = 0;
    for (int w = 0; w < kThreads / 64; ++w) s += part[w];
    bsum[blockIdx.x] = s;
  }
}

// exclusive scan of the block sums in place (one workgroup); total -> *total
__global__ void __launch_bounds__(kThreads)
k_scan_top(int nb, long long *__restrict__ bsum, long long *__restrict__ total) {
  __shared__ long long buf[kThreads];
  long long carry = 0;
  for (int b0 = 0; b0 < nb; b0 += kThreads) {
    const int b = b0 + threadIdx.x;
    const long long v = b < nb ? bsum[b] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kThreads; o <<= 1) {     // Hillis-Steele inclusive scan
      const long long add = threadIdx.x >= o ? buf[threadIdx.x - o] : 0;
      __syncthreads();
      buf[threadIdx.x] += add;
      __syncthreads();
    }
    if (b < nb) bsum[b] = carry + buf[threadIdx.x] - v;
    carry += buf[kThreads - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

// row(c) = base + exclusive prefix of (width + 1); the diagonal goes first in
// its row (Topology.f90:54-59)
__global__ void __launch_bounds__(kThreads)
k_scan_rows(int n, int base, const long long *__restrict__ bsum, const int *__restrict__ width,
            int *__restrict__ row, int *__restrict__ col) {
  __shared__ int wsum[kThreads / 64];
  const long long t0 = (long long)blockIdx.x * kScanTile;
  // a thread's kScanItems consecutive cells
  int v[kScanItems], acc = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const long long c = t0 + (long long)threadIdx.x * kScanItems + k;
    v[k] = c < n ? width[c] + 1 : 0;
    acc += v[k];
  }
  // exclusive scan of the per-thread totals: within the wave, then waves
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int inc = acc;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wv; ++w) before += wsum[w];
  long long off = bsum[blockIdx.x] + base + before + inc - acc;
  for (int k = 0; k < kScanItems; ++k) {
    const long long c = t0 + (long long)threadIdx.x * kScanItems + k;
    if (c < n) {
      row[c] = (int)off;
      col[off - base] = (int)c + base;
      if (c == n - 1) row[n] = (int)(off + v[k]);
    }
    off += v[k];
  }
}

// off-diagonal columns at the places k_width gave the face in its two rows
// (Topology.f90:62-71); the entry's tag 2 s + side locates pos(side + 1, s)
// after the sort; boundary faces get pos = base - 1 here
__global__ void __launch_bounds__(kThreads)
k_fill(long long nf, int base, const int *__restrict__ faces, const int *__restrict__ row,
       const int2 *__restrict__ slot, int *__restrict__ col, long long *__restrict__ tag,
       int *__restrict__ pos) {
  for (long long s = blockIdx.x * (long long)kThreads + threadIdx.x; s < nf;
       s += (long long)gridDim.x * kThreads) {
    const int2 f = reinterpret_cast<const int2 *>(faces)[s];
    if (f.y > 0) {
      const int c1 = f.x - 1, c2 = f.y - 1;
      const int2 k = slot[s];
      const long long q1 = row[c1] - base + k.x;
      const long long q2 = row[c2] - base + k.y;
      col[q1] = c2 + base;
      col[q2] = c1 + base;
      if (tag) {
        tag[q1] = 2 * s;
        tag[q2] = 2 * s + 1;
      }
    } else if (pos) {
      reinterpret_cast<int2 *>(pos)[s] = make_int2(base - 1, base - 1);
    }
  }
}

constexpr int kRegRow = 16;   // rows up to this long are sorted in registers

// One row of at most kRegRow entries: loaded once, sorted by an odd-even
// transposition network (stable: only strictly greater keys move down, so the
// padding keys stay behind every real entry), written back once.
__device__ __forceinline__ void sort_row_regs(int c, long long b, int w, int base,
                                              int *__restrict__ col, long long *__restrict__ tag,
                                              int *__restrict__ dia, int *__restrict__ pos) {
  int k[kRegRow];
  long long t[kRegRow];
#pragma unroll
  for (int i = 0; i < kRegRow; ++i) {
    k[i] = i < w ? col[b + i] : INT_MAX;
    t[i] = (tag && i < w) ? tag[b + i] : -1;
  }
#pragma unroll
  for (int pass = 0; pass < kRegRow; ++pass) {
#pragma unroll
    for (int i = pass & 1; i + 1 < kRegRow; i += 2) {
      if (k[i] > k[i + 1]) {
        const int kk = k[i];
        k[i] = k[i + 1];
        k[i + 1] = kk;
        const long long tt = t[i];
        t[i] = t[i + 1];
        t[i + 1] = tt;
      }
    }
  }
  int d = 0;
#pragma unroll
  for (int i = 0; i < kRegRow; ++i) {
    if (i < w) {
      col[b + i] = k[i];
      if (k[i] == c + base) d = i;
    }
  }
  dia[c] = (int)(b + d + base);
  if (!tag) return;
  // from the end: `last` is the end of the current run of equal columns
  int last = kRegRow - 1;
#pragma unroll
  for (int i = kRegRow - 1; i >= 0; --i) {
    if (i < w) {
      if (i == w - 1 || k[i] != k[last]) last = i;
      if (t[i] >= 0) pos[t[i]] = (int)(b + last + base);
    }
  }
}

// ascending columns per row with their tags, the diagonal's position
// (Topology.f90:75-82), and the face positions (:106-121)
__global__ void __launch_bounds__(kThreads)
k_sort(int n, int base, const int *__restrict__ row, int *__restrict__ col,
       long long *__restrict__ tag, int *__restrict__ dia, int *__restrict__ pos) {
  for (int c = blockIdx.x * kThreads + threadIdx.x; c < n; c += gridDim.x * kThreads) {
    const long long b = row[c] - base, e = row[c + 1] - base;
    if (e - b <= kRegRow) {
      sort_row_regs(c, b, (int)(e - b), base, col, tag, dia, pos);
      continue;
    }
    // longer rows: insertion sort in place
    for (long long q = b + 1; q < e; ++q) {
      const int key = col[q];
      const long long t = tag ? tag[q] : 0;
      long long r = q;
      while (r > b && col[r - 1] > key) {
        col[r] = col[r - 1];
        if (tag) tag[r] = tag[r - 1];
        --r;
      }
      col[r] = key;
      if (tag) tag[r] = t;
    }
    long long d = b;
    for (long long q = b; q < e; ++q)
      if (col[q] == c + base) d = q;
    dia[c] = (int)(d + base);
    if (!tag) continue;
    // walk down the row: `last` is the end of the current run of equal columns
    long long last = e - 1;
    for (long long q = e - 1; q >= b; --q) {
      if (col[q] != col[last]) last = q;
      const long long t = tag[q];
      if (t >= 0) pos[t] = (int)(last + base);
    }
  }
}

}  // namespace

extern "C" int tfs_topology_build_device(int n_cells, long long n_faces, const int *d_faces_c,
                                         int base, long long col_cap, int *d_row, int *d_col,
                                         int *d_dia, int *d_pos, long long *nnz) {
  if (n_cells < 1 || n_faces < 0 || (base != 0 && base != 1) || !d_row || !d_col || !d_dia ||
      (n_faces > 0 && !d_faces_c) || col_cap < n_cells)
    return TFS_ERR_INVALID;
  // faces and pos are read and written as (c1, c2) / (pos1, pos2) pairs
  if ((reinterpret_cast<uintptr_t>(d_faces_c) | reinterpret_cast<uintptr_t>(d_pos)) % 8 != 0)
    return TFS_ERR_INVALID;
  if (nnz) *nnz = 0;
  const int nb = (int)((n_cells + (long long)kScanTile - 1) / kScanTile);
  int rc = TFS_OK;
  int *width = nullptr, *err = nullptr;
  int2 *slot = nullptr;
  long long *bsum = nullptr, *total = nullptr, *tag = nullptr;
  int h_err = 0;
  long long h_total = 0;
  const hipStream_t st = nullptr;    // the calling thread's default stream
  // scratch from the stream-ordered pool: no device-wide synchronisation on
  // allocation or release
  if (hipMallocAsync((void **)&width, sizeof(int) * (size_t)n_cells, st) != hipSuccess ||
      hipMallocAsync((void **)&bsum, sizeof(long long) * (size_t)nb + 2 * sizeof(long long), st) !=
          hipSuccess ||
      hipMallocAsync((void **)&slot, sizeof(int2) * (size_t)(n_faces > 0 ? n_faces : 1), st) !=
          hipSuccess) {
    (void)hipGetLastError();
    rc = TFS_ERR_NOMEM;
    goto done;
  }
  total = bsum + nb;
  err = reinterpret_cast<int *>(bsum + nb + 1);
  hipLaunchKernelGGL(k_fill_int, dim3(grid_of(n_cells)), dim3(kThreads), 0, st,
                     (long long)n_cells, 0, width);
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(kThreads), 0, st, 1LL, 0, err);
  if (n_faces > 0)
    hipLaunchKernelGGL(k_width, dim3(grid_of(n_faces)), dim3(kThreads), 0, st, n_cells, n_faces,
                       d_faces_c, width, slot, err);
  hipLaunchKernelGGL(k_scan_sums, dim3(nb), dim3(kThreads), 0, st, n_cells, width, bsum);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kThreads), 0, st, nb, bsum, total);
  TFS_T_HIP(hipGetLastError());
  TFS_T_HIP(hipMemcpy(&h_total, total, sizeof(long long), hipMemcpyDeviceToHost));
  TFS_T_HIP(hipMemcpy(&h_err, err, sizeof(int), hipMemcpyDeviceToHost));
  if (h_err) {
    rc = TFS_ERR_INVALID;
    goto done;
  }
  if (h_total > INT_MAX - 1 || h_total > col_cap) {
    rc = TFS_ERR_INVALID;
    if (nnz) *nnz = h_total;
    goto done;
  }
  if (d_pos) {
    if (hipMallocAsync((void **)&tag, sizeof(long long) * (size_t)h_total, st) != hipSuccess) {
      (void)hipGetLastError();
      rc = TFS_ERR_NOMEM;
      goto done;
    }
    // the diagonal entries keep tag -1 (k_scan_rows writes only their column)
    hipLaunchKernelGGL(k_fill_int, dim3(grid_of(2 * h_total)), dim3(kThreads), 0, st,
                       2 * h_total, -1, reinterpret_cast<int *>(tag));
  }
  hipLaunchKernelGGL(k_scan_rows, dim3(nb), dim3(kThreads), 0, st, n_cells, base, bsum, width,
                     d_row, d_col);
  if (n_faces > 0)
    hipLaunchKernelGGL(k_fill, dim3(grid_of(n_faces)), dim3(kThreads), 0, st, n_faces, base,
                       d_faces_c, d_row, slot, d_col, tag, d_pos);
  hipLaunchKernelGGL(k_sort, dim3(grid_of(n_cells)), dim3(kThreads), 0, st, n_cells, base, d_row,
                     d_col, tag, d_dia, d_pos);
  TFS_T_HIP(hipGetLastError());
  TFS_T_HIP(hipStreamSynchronize(st));
  if (nnz) *nnz = h_total;
done:
  for (void *p : {(void *)width, (void *)bsum, (void *)slot, (void *)tag})
    if (p) (void)hipFreeAsync(p, st);
  return rc;
}
