// tfs_device.hpp -- device-side building blocks for gfx950 (wave64, CDNA4).
//
// Cross-lane sums use DPP (quad permutes and row rotations inside each row of
// 16 lanes) and two lane swaps across rows; block and grid reductions are
// deterministic: fixed trees, and the grid stage runs in whichever workgroup
// finishes last, over the per-workgroup partials in index order.
#pragma once

#include <hip/hip_runtime.h>

namespace tfs {

constexpr int kWave = 64;
constexpr double kTiny = 1.e-30;  // Sources/Shared/allp_mod.f90:40

// Scalars of one Krylov solve, resident on the device so that no kernel needs
// a host round trip; every kernel of an iteration reads `done` first and
// returns at once after convergence.
struct KState {
  double rho, rho_old, alpha, beta, pq, err, bnrm2, tol, ini_res, fin_res;
  int iter, cap, done, niter_out, status, pad;
  unsigned int ctr[4];
  double aux[6];  // dot products and norms kept for the host (no state transition)
};

// status values mirrored from the public header
constexpr int kStOk = 0, kStNotConverged = 1, kStBreakdown = -3;

template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int l2 = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, false);
  const int h2 = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(h2, l2);
}

template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over the 64 lanes; the value of lane 0 is the result (other lanes hold
// the same sum with a different association).
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]  (lane ^ 1)
  v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]  (lane ^ 2)
  v += dpp_mov<0x124>(v);  // row_ror:4
  v += dpp_mov<0x128>(v);  // row_ror:8  -> every lane holds its row's sum
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  return v;
}

// Block sum for NT threads; result valid in thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *lds) {
  v = wave_sum(v);
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  if (l == 0) lds[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x < kWave) {
    t = l < NT / kWave ? lds[l] : 0.0;
    t = wave_sum(t);
  }
  __syncthreads();
  return t;
}

__device__ __forceinline__ bool solve_done(const KState *st) {
  return __builtin_amdgcn_readfirstlane(st->done) != 0;
}

}  // namespace tfs
