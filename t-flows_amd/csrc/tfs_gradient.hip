// tfs_gradient.hip -- least-squares cell gradients on the device.
//
// T-FlowS computes cell gradients by least squares over the face neighbours:
//   Compute_Gradient_Matrix  Sources/Process/Compute_Gradient_Matrix.f90:21-116
//     G(c) = inverse of sum over the faces of c of d d^T (d the centre-to-centre
//     vector of the face after Compute_Geometry), by cofactors over Jac + TINY;
//   GraPhi                   Sources/Process/GraPhi.f90:25-173
//     grad phi(c) = G(c) sum over the faces of c of (phi(c2) - phi(c1)) d,
//     zero difference across a SYMMETRY boundary;
//   GradP                    Sources/Process/GradP.f90:24-64
//     boundary cells (not BUFFER / PRESSURE) take the inside value, gradient,
//     boundary values extrapolated along it with factor 1.2, gradient again.
// All with Boundary = .TRUE., as every call site passes it: both cells of an
// interior face and the inside cell of a boundary face take the face's term.
//
// The reference accumulates in one sequential loop over faces.  Here every
// cell sums its own faces -- kept in ascending face order in a per-cell
// incidence list built once -- so each sum has the reference's order and the
// result is bitwise the same, one thread per cell and no atomics.  The three
// components of a gradient are formed in one pass (the reference makes three
// passes over the same data, one per component; each component's sum is
// unchanged).  The static mesh data of every incidence (d, the other cell, the
// side) is laid out in the cells' incidence order once, at creation, so the
// per-call traffic is contiguous except for one gathered phi per incidence;
// per cell, G, phi and the three components.
//
// The same mesh object serves Correct_Velocity
// (Sources/Process/Correct_Velocity.f90:34-135): velocities corrected by the
// pressure gradient, PRESSURE boundary cells given the inside velocity, face
// fluxes corrected by the pressure-correction difference, the mass error per
// cell (each cell sums its faces' fluxes in face order, as the reference's
// loop does) and the maxima of |mass error|, CFL and Peclet numbers (maxima
// are exact in any order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "tfs_solver.h"

namespace {

constexpr int kThreads = 256;
constexpr double kTiny = 1.e-30;   // Sources/Shared/allp_mod.f90:40

#define TFS_G_HIP(expr)                        \
  do {                                         \
    if ((expr) != hipSuccess) {                \
      (void)hipGetLastError();                 \
      return TFS_ERR_DEVICE;                   \
    }                                          \
  } while (0)

int grid_of(long long m) {
  long long g = (m + kThreads - 1) / kThreads;
  return (int)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

// G(c): the normal matrix summed over the cell's faces in face order, then its
// inverse (Compute_Gradient_Matrix.f90:30-70, :96-116).  id = d of every
// incidence (three blocks of m); g holds six blocks of n: G(1..6, c) at
// g[k n + c].
__global__ void __launch_bounds__(kThreads)
k_gmatrix(int n, long long m, const int *__restrict__ ptr, const double *__restrict__ id,
          double *__restrict__ g) {
  for (int c = blockIdx.x * kThreads + threadIdx.x; c < n; c += gridDim.x * kThreads) {
    double G0 = 0.0, G1 = 0.0, G2 = 0.0, G3 = 0.0, G4 = 0.0, G5 = 0.0;
    for (int k = ptr[c]; k < ptr[c + 1]; ++k) {
      const double x = id[k], y = id[m + k], z = id[2 * m + k];
      G0 = G0 + x * x;
      G1 = G1 + y * y;
      G2 = G2 + z * z;
      G3 = G3 + x * y;
      G4 = G4 + x * z;
      G5 = G5 + y * z;
    }
    const double jac = G0 * G1 * G2 - G0 * G5 * G5 - G3 * G3 * G2 + 2.0 * G3 * G4 * G5 -
                       G4 * G4 * G1;
    g[c] = +(G1 * G2 - G5 * G5) / (jac + kTiny);
    g[n + c] = +(G0 * G2 - G4 * G4) / (jac + kTiny);
    g[2LL * n + c] = +(G0 * G1 - G3 * G3) / (jac + kTiny);
    g[3LL * n + c] = -(G3 * G2 - G4 * G5) / (jac + kTiny);
    g[4LL * n + c] = +(G3 * G5 - G4 * G1) / (jac + kTiny);
    g[5LL * n + c] = -(G0 * G5 - G3 * G4) / (jac + kTiny);
  }
}

// GraPhi for i = 1, 2, 3 at once (GraPhi.f90:25-73, :75-123, :125-173 with
// Boundary = .TRUE.).  The mesh data of every incidence -- the face vector d
// (three blocks of m), the phi slot of the face's other cell, and flags
// (bit 0: this cell is the face's c2; bit 1: a SYMMETRY boundary face) -- sit
// in the cells' incidence order, so a wavefront reads them in contiguous
// runs; only phi of the other cell is gathered.  phi(c2) - phi(c1) is formed
// from the same two values whichever side evaluates it.
__global__ void __launch_bounds__(kThreads)
k_graphi(int n, long long m, const int *__restrict__ ptr, const int *__restrict__ other,
         const unsigned char *__restrict__ flag, const double *__restrict__ id,
         const double *__restrict__ g, const double *__restrict__ phi, double *__restrict__ px,
         double *__restrict__ py, double *__restrict__ pz) {
  for (int c = blockIdx.x * kThreads + threadIdx.x; c < n; c += gridDim.x * kThreads) {
    const double g0 = g[c], g1 = g[n + c], g2 = g[2LL * n + c], g3 = g[3LL * n + c],
                 g4 = g[4LL * n + c], g5 = g[5LL * n + c];
    const double pc = phi[c];
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int k = ptr[c]; k < ptr[c + 1]; ++k) {
      const double po = phi[other[k]];
      const unsigned fl = flag[k];
      double dphi = (fl & 1u) ? pc - po : po - pc;
      if (fl & 2u) dphi = 0.0;
      const double x = id[k], y = id[m + k], z = id[2 * m + k];
      ax = ax + dphi * (g0 * x + g3 * y + g4 * z);
      ay = ay + dphi * (g3 * x + g1 * y + g5 * z);
      az = az + dphi * (g4 * x + g5 * y + g2 * z);
    }
    px[c] = ax;
    py[c] = ay;
    pz[c] = az;
  }
}

// GradP.f90:32-40: the boundary cells listed in bf (inside cell, boundary slot)
// take the inside value
__global__ void __launch_bounds__(kThreads)
k_bnd_copy(long long m, const int2 *__restrict__ bf, double *__restrict__ phi) {
  for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < m;
       i += (long long)gridDim.x * kThreads) {
    const int2 e = bf[i];
    phi[e.y] = phi[e.x];
  }
}

// GradP.f90:48-58: extrapolation along the inside cell's gradient
__global__ void __launch_bounds__(kThreads)
k_bnd_extrap(long long m, long long na, const int2 *__restrict__ bf, const double *__restrict__ xc,
             const double *__restrict__ px, const double *__restrict__ py,
             const double *__restrict__ pz, double *__restrict__ phi) {
  for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < m;
       i += (long long)gridDim.x * kThreads) {
    const int2 e = bf[i];
    const int c1 = e.x, c2 = e.y;
    phi[c2] = phi[c1] + 1.2 * (px[c1] * (xc[c2] - xc[c1]) + py[c1] * (xc[na + c2] - xc[na + c1]) +
                               pz[c1] * (xc[2 * na + c2] - xc[2 * na + c1]));
  }
}

// ---- Correct_Velocity -------------------------------------------------------

constexpr int kParts = 8192;   // the largest grid of grid_of(): one partial per block

// max over the workgroup of v, written by thread 0 to part[blockIdx.x]
__device__ void block_max(double v, double *__restrict__ part) {
  __shared__ double wmax[kThreads / 64];
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wmax[0];
    for (int w = 1; w < kThreads / 64; ++w) m = fmax(m, wmax[w]);
    part[blockIdx.x] = m;
  }
  __syncthreads();
}

// Correct_Velocity.f90:34-46
__global__ void __launch_bounds__(kThreads)
k_cv_cells(int n, const double *__restrict__ vol, const double *__restrict__ sav,
           const double *__restrict__ px, const double *__restrict__ py,
           const double *__restrict__ pz, double *__restrict__ u, double *__restrict__ v,
           double *__restrict__ w) {
  for (int c = blockIdx.x * kThreads + threadIdx.x; c < n; c += gridDim.x * kThreads) {
    const double vc = vol[c], sc = sav[c];
    u[c] = u[c] - px[c] * vc / sc;
    v[c] = v[c] - py[c] * vc / sc;
    w[c] = w[c] - pz[c] * vc / sc;
  }
}

// Correct_Velocity.f90:48-59 (PRESSURE boundaries), :72-82 (flux correction),
// :117-135 (CFL and Peclet numbers of the corrected fluxes); per-block maxima
// to part_cfl / part_pe
__global__ void __launch_bounds__(kThreads)
k_cv_faces(int n, long long nf, const int2 *__restrict__ faces, const unsigned char *__restrict__ kind,
           const double *__restrict__ pp, const double *__restrict__ coef,
           const double *__restrict__ bou, const unsigned char *__restrict__ in_m,
           const double *__restrict__ scoef, const double *__restrict__ d, double dt,
           double visc, double *__restrict__ flux, double *__restrict__ u,
           double *__restrict__ v, double *__restrict__ w, double *__restrict__ part_cfl,
           double *__restrict__ part_pe) {
  double cm = 0.0, pm = 0.0;
  for (long long s = blockIdx.x * (long long)kThreads + threadIdx.x; s < nf;
       s += (long long)gridDim.x * kThreads) {
    const int2 f = faces[s];
    const int k1 = f.x - 1;
    const int kb = f.y < 0 ? -f.y - 1 : 0;
    const int kd = f.y < 0 ? kind[kb] : TFS_BC_WALL;
    if (f.y < 0 && kd == TFS_BC_PRESSURE) {
      const long long k2 = (long long)n + kb;
      u[k2] = u[k1];
      v[k2] = v[k1];
      w[k2] = w[k1];
    }
    if (f.y > 0 || kd == TFS_BC_BUFFER) {
      const long long k2 = f.y > 0 ? (long long)(f.y - 1) : (long long)n + kb;
      const double a = f.y > 0 ? coef[s] : (bou ? bou[kb] : 0.0);
      const double fl = flux[s] + (pp[k2] - pp[k1]) * a;
      flux[s] = fl;
      if (!in_m || in_m[s]) {
        const double x = d[s], y = d[nf + s], z = d[2 * nf + s];
        cm = fmax(cm, fabs(dt * fl / (scoef[s] * (x * x + y * y + z * z))));
        pm = fmax(pm, fabs(fl / scoef[s] / (visc + kTiny)));
      }
    }
  }
  block_max(cm, part_cfl);
  block_max(pm, part_pe);
}

// Correct_Velocity.f90:88-111: b(c) = -sum of outgoing fluxes + incoming, in
// face order, over vol * density; per-block maxima of |b| to part_err
__global__ void __launch_bounds__(kThreads)
k_cv_mass(int n, const int *__restrict__ ptr, const long long *__restrict__ inc,
          const double *__restrict__ flux, const double *__restrict__ vol,
          const double *__restrict__ den, double *__restrict__ b, double *__restrict__ part_err) {
  double em = 0.0;
  for (int c = blockIdx.x * kThreads + threadIdx.x; c < n; c += gridDim.x * kThreads) {
    double m = 0.0;
    for (int k = ptr[c]; k < ptr[c + 1]; ++k) {
      const long long e = inc[k];
      const double fl = flux[e >> 1];
      m = (e & 1) ? m + fl : m - fl;
    }
    m = m / (vol[c] * den[c]);
    b[c] = m;
    em = fmax(em, fabs(m));
  }
  block_max(em, part_err);
}

// the three maxima over the per-block partials
__global__ void __launch_bounds__(kThreads)
k_cv_final(int nf_parts, int n_parts, const double *__restrict__ part, double *__restrict__ out) {
  double e = 0.0, cm = 0.0, pm = 0.0;
  for (int i = threadIdx.x; i < n_parts; i += kThreads) e = fmax(e, part[i]);
  for (int i = threadIdx.x; i < nf_parts; i += kThreads) {
    cm = fmax(cm, part[kParts + i]);
    pm = fmax(pm, part[2 * kParts + i]);
  }
  block_max(e, out);
  block_max(cm, out + 1);
  block_max(pm, out + 2);
}

// ---- Compute_Pressure_Fractional's face loop -----------------------------------

// Per cell slot, once per call, what the faces gather: the velocity with the
// pressure-gradient term put back, u + p_x vol / sav (the raw velocity for a
// boundary cell, which the reference's boundary faces read), sav, p and the
// density -- six values side by side, so a face reads two 48-byte records
// instead of twenty scattered arrays.  Each value is the one the reference's
// face expression forms, so the result stays bitwise.
__global__ void __launch_bounds__(kThreads)
k_pf_cells(int n, long long na, const double *__restrict__ vol, const double *__restrict__ sav,
           const double *__restrict__ u, const double *__restrict__ v,
           const double *__restrict__ w, const double *__restrict__ px,
           const double *__restrict__ py, const double *__restrict__ pz,
           const double *__restrict__ p, const double *__restrict__ den,
           double *__restrict__ rec) {
  for (long long c = blockIdx.x * (long long)kThreads + threadIdx.x; c < na;
       c += (long long)gridDim.x * kThreads) {
    double *r = rec + 6 * c;
    if (c < n) {
      r[0] = u[c] + px[c] * vol[c] / sav[c];
      r[1] = v[c] + py[c] * vol[c] / sav[c];
      r[2] = w[c] + pz[c] * vol[c] / sav[c];
      r[3] = sav[c];
    } else {
      r[0] = u[c];
      r[1] = v[c];
      r[2] = w[c];
      r[3] = 0.0;
    }
    r[4] = p[c];
    r[5] = den[c];
  }
}

// Compute_Pressure_Fractional.f90:50-131 for fluid materials without BUFFER
// faces: per face the f-weighted density, the face velocity rebuilt from the
// cells' velocities and pressure gradients (Rhie-Chow), the flux and A12;
// a boundary face takes its boundary cell's velocity and has A12 = 0.
__global__ void __launch_bounds__(kThreads)
k_pf_faces(int n, long long nf, const int2 *__restrict__ faces, const double *__restrict__ f,
           const double *__restrict__ sv, const double *__restrict__ rec,
           double *__restrict__ flux, double *__restrict__ a12) {
  for (long long s = blockIdx.x * (long long)kThreads + threadIdx.x; s < nf;
       s += (long long)gridDim.x * kThreads) {
    const int2 fc = faces[s];
    const long long k1 = fc.x - 1;
    const long long k2 = fc.y > 0 ? (long long)(fc.y - 1) : (long long)n + (-fc.y - 1);
    const double *r1 = rec + 6 * k1, *r2 = rec + 6 * k2;
    const double fs = f[s], sx = sv[s], sy = sv[nf + s], sz = sv[2 * nf + s];
    const double dens = fs * r1[5] + (1.0 - fs) * r2[5];
    if (fc.y > 0) {
      const double sav1 = r1[3], sav2 = r2[3];
      double us = fs * r1[0] + (1.0 - fs) * r2[0];
      double vs = fs * r1[1] + (1.0 - fs) * r2[1];
      double ws = fs * r1[2] + (1.0 - fs) * r2[2];
      const double wsav = fs / sav1 + (1.0 - fs) / sav2;
      const double dp = r1[4] - r2[4];
      us = us + dp * sx * wsav;
      vs = vs + dp * sy * wsav;
      ws = ws + dp * sz * wsav;
      flux[s] = dens * (us * sx + vs * sy + ws * sz);
      const double a = dens * (sx * sx + sy * sy + sz * sz);
      a12[s] = a * wsav;
    } else {
      flux[s] = dens * (r2[0] * sx + r2[1] * sy + r2[2] * sz);
      a12[s] = 0.0;
    }
  }
}

// Compute_Pressure_Simple.f90:50-177 for fluid materials without BUFFER
// faces: per interior face SMDPN = |S|^2 / (S . d), A12 = 0.5 DENs SMDPN
// (vol1/sav1 + vol2/sav2), the interpolated velocity and pressure gradient and
// the face flux; per boundary face by the kind of its boundary cell: the flux
// from the boundary cell's velocity (inflow, outflow, convective) or the
// inside cell's (pressure), A12 = DENs SMDPN vol1/sav1 added to the diagonal
// (bnd_on) for outflow, convective and pressure boundaries, no flux on walls
// and symmetry planes.  Every expression in the reference's order.
__global__ void __launch_bounds__(kThreads)
k_ps_faces(int n, long long nf, const int2 *__restrict__ faces,
           const unsigned char *__restrict__ kind, const double *__restrict__ f,
           const double *__restrict__ sv, const double *__restrict__ dv,
           const double *__restrict__ vol, const double *__restrict__ sav,
           const double *__restrict__ u, const double *__restrict__ v,
           const double *__restrict__ w, const double *__restrict__ p,
           const double *__restrict__ px, const double *__restrict__ py,
           const double *__restrict__ pz, const double *__restrict__ den,
           double *__restrict__ flux, double *__restrict__ a12,
           unsigned char *__restrict__ bnd_on) {
  for (long long s = blockIdx.x * (long long)kThreads + threadIdx.x; s < nf;
       s += (long long)gridDim.x * kThreads) {
    const int2 fc = faces[s];
    const long long k1 = fc.x - 1;
    const long long k2 = fc.y > 0 ? (long long)(fc.y - 1) : (long long)n + (-fc.y - 1);
    const double fs = f[s], sx = sv[s], sy = sv[nf + s], sz = sv[2 * nf + s];
    const double dx = dv[s], dy = dv[nf + s], dz = dv[2 * nf + s];
    const double dens = fs * den[k1] + (1.0 - fs) * den[k2];
    double fl = 0.0, a = 0.0;
    unsigned char on = 0;
    if (fc.y > 0) {
      const double smdpn = (sx * sx + sy * sy + sz * sz) / (sx * dx + sy * dy + sz * dz);
      const double us = fs * u[k1] + (1.0 - fs) * u[k2];
      const double vs = fs * v[k1] + (1.0 - fs) * v[k2];
      const double ws = fs * w[k1] + (1.0 - fs) * w[k2];
      a = 0.5 * dens * smdpn * (vol[k1] / sav[k1] + vol[k2] / sav[k2]);
      const double dpxi = .5 * (px[k1] + px[k2]) * dx;
      const double dpyi = .5 * (py[k1] + py[k2]) * dy;
      const double dpzi = .5 * (pz[k1] + pz[k2]) * dz;
      fl = dens * (us * sx + vs * sy + ws * sz) + a * (p[k1] - p[k2]) + a * (dpxi + dpyi + dpzi);
    } else {
      const int kd = kind[-fc.y - 1];
      if (kd == TFS_BC_INFLOW || kd == TFS_BC_OUTFLOW || kd == TFS_BC_CONVECT ||
          kd == TFS_BC_PRESSURE) {
        const long long kv = kd == TFS_BC_PRESSURE ? k1 : k2;
        fl = dens * (u[kv] * sx + v[kv] * sy + w[kv] * sz);
        if (kd != TFS_BC_INFLOW) {
          const double smdpn = (sx * sx + sy * sy + sz * sz) / (sx * dx + sy * dy + sz * dz);
          a = dens * smdpn * vol[k1] / sav[k1];
          on = 1;
        }
      }
    }
    flux[s] = fl;
    a12[s] = a;
    bnd_on[s] = on;
  }
}

// b(c) = inflow - outflow over the cell's faces in face order
// (Compute_Pressure_Fractional.f90:112-113, :126)
__global__ void __launch_bounds__(kThreads)
k_pf_rhs(int n, const int *__restrict__ ptr, const long long *__restrict__ inc,
         const double *__restrict__ flux, double *__restrict__ b) {
  for (int c = blockIdx.x * kThreads + threadIdx.x; c < n; c += gridDim.x * kThreads) {
    double m = 0.0;
    for (int k = ptr[c]; k < ptr[c + 1]; ++k) {
      const long long e = inc[k];
      const double fl = flux[e >> 1];
      m = (e & 1) ? m + fl : m - fl;
    }
    b[c] = m;
  }
}

}  // namespace

struct tfs_gradient {
  int n = 0, nb = 0;
  long long nf = 0, nbf = 0;
  int *ptr = nullptr;
  long long *inc = nullptr;
  int2 *faces = nullptr, *bf = nullptr;
  long long m = 0;            // incidences (interior faces twice, boundary faces once)
  bool has_buffer = false;    // some boundary cell is of kind TFS_BC_BUFFER
  double *d = nullptr, *g = nullptr, *xc = nullptr, *id = nullptr;
  int *other = nullptr;
  unsigned char *kind = nullptr, *flag = nullptr;
  double *part = nullptr;     // 3 kParts partial maxima + 3 results
  double *rec = nullptr;      // 6 values per cell slot for the pressure face loop (lazy)
};

extern "C" int tfs_gradient_destroy(tfs_gradient *S) {
  if (!S) return TFS_OK;
  for (void *p : {(void *)S->ptr, (void *)S->inc, (void *)S->faces, (void *)S->bf, (void *)S->d,
                  (void *)S->g, (void *)S->xc, (void *)S->kind,
                  (void *)S->part, (void *)S->id, (void *)S->other, (void *)S->flag,
                  (void *)S->rec})
    if (p) (void)hipFree(p);
  delete S;
  return TFS_OK;
}

extern "C" int tfs_gradient_create(int n_cells, int n_bnd_cells, long long n_faces,
                                   const int *faces_c, const double *d, const double *xc,
                                   const unsigned char *bc, tfs_gradient **out) {
  if (!out) return TFS_ERR_INVALID;
  *out = nullptr;
  if (n_cells < 1 || n_bnd_cells < 0 || n_faces < 0 || (n_faces > 0 && (!faces_c || !d)))
    return TFS_ERR_INVALID;
  const int n = n_cells, nb = n_bnd_cells;
  const long long nf = n_faces;
  std::vector<int> cnt((size_t)n + 1, 0), seen((size_t)nb, 0);
  std::vector<unsigned char> sym((size_t)std::max(nb, 1), 0), kinds((size_t)std::max(nb, 1), 0);
  std::vector<int2> bf;
  bool any_buffer = false;
  for (long long s = 0; s < nf; ++s) {
    const int c1 = faces_c[2 * s], c2 = faces_c[2 * s + 1];
    if (c1 < 1 || c1 > n || c2 == 0 || c2 > n || c2 < -nb) return TFS_ERR_INVALID;
    cnt[(size_t)c1]++;
    if (c2 > 0) {
      cnt[(size_t)c2]++;
    } else {
      // a boundary cell belongs to one face: GradP writes it from that face
      if (seen[(size_t)(-c2 - 1)]++) return TFS_ERR_INVALID;
      const int kind = bc ? bc[-c2 - 1] : TFS_BC_WALL;
      if (kind > TFS_BC_CONVECT) return TFS_ERR_INVALID;
      kinds[(size_t)(-c2 - 1)] = (unsigned char)kind;
      any_buffer = any_buffer || kind == TFS_BC_BUFFER;
      if (kind == TFS_BC_SYMMETRY) sym[(size_t)(-c2 - 1)] = 1;
      if (kind != TFS_BC_BUFFER && kind != TFS_BC_PRESSURE)
        bf.push_back(make_int2(c1 - 1, n + (-c2 - 1)));
    }
  }
  // cell -> incident faces in ascending face order (a counting sort over the
  // faces in order keeps each list sorted)
  for (int c = 0; c < n; ++c) cnt[(size_t)c + 1] += cnt[(size_t)c];
  std::vector<long long> inc((size_t)cnt[(size_t)n]);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (long long s = 0; s < nf; ++s) {
    const int c1 = faces_c[2 * s] - 1, c2 = faces_c[2 * s + 1];
    inc[(size_t)fill[(size_t)c1]++] = 2 * s;
    if (c2 > 0) inc[(size_t)fill[(size_t)(c2 - 1)]++] = 2 * s + 1;
  }
  // per incidence, in the same order: d of the face, the phi slot of the
  // face's other cell, side and symmetry flags
  const size_t m = inc.size();
  std::vector<double> id(3 * m);
  std::vector<int> other(m);
  std::vector<unsigned char> flag(m);
  for (size_t k = 0; k < m; ++k) {
    const long long e = inc[k], s = e >> 1;
    const int c1 = faces_c[2 * s], c2 = faces_c[2 * s + 1];
    for (int a = 0; a < 3; ++a) id[(size_t)a * m + k] = d[(size_t)a * (size_t)nf + (size_t)s];
    if (e & 1) {
      other[k] = c1 - 1;
      flag[k] = 1;
    } else {
      other[k] = c2 > 0 ? c2 - 1 : n + (-c2 - 1);
      flag[k] = (c2 < 0 && sym[(size_t)(-c2 - 1)]) ? 2 : 0;
    }
  }
  auto *S = new tfs_gradient;
  S->n = n;
  S->nb = nb;
  S->nf = nf;
  S->m = (long long)m;
  S->has_buffer = any_buffer;
  S->nbf = (long long)bf.size();
  auto up = [](void **dst, const void *src, size_t bytes) -> bool {
    if (hipMalloc(dst, std::max<size_t>(bytes, 8)) != hipSuccess) return false;
    return bytes == 0 || !src || hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  const size_t na = (size_t)n + (size_t)nb;
  bool ok = up((void **)&S->ptr, cnt.data(), sizeof(int) * cnt.size()) &&
            up((void **)&S->inc, inc.data(), sizeof(long long) * inc.size()) &&
            up((void **)&S->faces, faces_c, sizeof(int2) * (size_t)nf) &&
            up((void **)&S->bf, bf.data(), sizeof(int2) * bf.size()) &&
            up((void **)&S->d, d, 3 * sizeof(double) * (size_t)nf) &&
            up((void **)&S->g, nullptr, 6 * sizeof(double) * (size_t)n) &&
            up((void **)&S->kind, kinds.data(), kinds.size()) &&
            up((void **)&S->part, nullptr, sizeof(double) * (3 * (size_t)kParts + 3)) &&
            up((void **)&S->id, id.data(), sizeof(double) * id.size()) &&
            up((void **)&S->other, other.data(), sizeof(int) * other.size()) &&
            up((void **)&S->flag, flag.data(), flag.size()) &&
            (!xc || up((void **)&S->xc, xc, 3 * sizeof(double) * na));
  if (!ok) {
    (void)hipGetLastError();
    tfs_gradient_destroy(S);
    return TFS_ERR_NOMEM;
  }
  hipLaunchKernelGGL(k_gmatrix, dim3(grid_of(n)), dim3(kThreads), 0, nullptr, n, S->m, S->ptr,
                     S->id, S->g);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(nullptr) != hipSuccess) {
    (void)hipGetLastError();
    tfs_gradient_destroy(S);
    return TFS_ERR_DEVICE;
  }
  *out = S;
  return TFS_OK;
}

extern "C" int tfs_gradient_matrix_get(tfs_gradient *S, double *g) {
  if (!S || !g) return TFS_ERR_INVALID;
  std::vector<double> soa(6 * (size_t)S->n);
  TFS_G_HIP(hipMemcpy(soa.data(), S->g, sizeof(double) * soa.size(), hipMemcpyDeviceToHost));
  for (int c = 0; c < S->n; ++c)
    for (int k = 0; k < 6; ++k) g[6 * (size_t)c + k] = soa[(size_t)k * S->n + c];
  return TFS_OK;
}

extern "C" int tfs_grad_phi_device(tfs_gradient *S, const double *d_phi, double *d_px,
                                   double *d_py, double *d_pz) {
  if (!S || !d_phi || !d_px || !d_py || !d_pz) return TFS_ERR_INVALID;
  hipLaunchKernelGGL(k_graphi, dim3(grid_of(S->n)), dim3(kThreads), 0, nullptr, S->n, S->m,
                     S->ptr, S->other, S->flag, S->id, S->g, d_phi, d_px, d_py, d_pz);
  TFS_G_HIP(hipGetLastError());
  TFS_G_HIP(hipStreamSynchronize(nullptr));
  return TFS_OK;
}

extern "C" int tfs_grad_p_device(tfs_gradient *S, double *d_phi, double *d_px, double *d_py,
                                 double *d_pz) {
  if (!S || !d_phi || !d_px || !d_py || !d_pz || (S->nbf > 0 && !S->xc)) return TFS_ERR_INVALID;
  const long long na = (long long)S->n + S->nb;
  if (S->nbf > 0)
    hipLaunchKernelGGL(k_bnd_copy, dim3(grid_of(S->nbf)), dim3(kThreads), 0, nullptr, S->nbf,
                       S->bf, d_phi);
  hipLaunchKernelGGL(k_graphi, dim3(grid_of(S->n)), dim3(kThreads), 0, nullptr, S->n, S->m,
                     S->ptr, S->other, S->flag, S->id, S->g, d_phi, d_px, d_py, d_pz);
  if (S->nbf > 0)
    hipLaunchKernelGGL(k_bnd_extrap, dim3(grid_of(S->nbf)), dim3(kThreads), 0, nullptr, S->nbf,
                       na, S->bf, S->xc, d_px, d_py, d_pz, d_phi);
  hipLaunchKernelGGL(k_graphi, dim3(grid_of(S->n)), dim3(kThreads), 0, nullptr, S->n, S->m,
                     S->ptr, S->other, S->flag, S->id, S->g, d_phi, d_px, d_py, d_pz);
  TFS_G_HIP(hipGetLastError());
  TFS_G_HIP(hipStreamSynchronize(nullptr));
  return TFS_OK;
}

extern "C" int tfs_correct_velocity_device(tfs_gradient *S, const double *d_vol,
                                           const double *d_sav, const double *d_px,
                                           const double *d_py, const double *d_pz, double *d_u,
                                           double *d_v, double *d_w, const double *d_pp,
                                           const double *d_coef, const double *d_bou,
                                           const double *d_den, const unsigned char *d_in_m,
                                           const double *d_scoef, double dt, double visc,
                                           double *d_flux, double *d_b, double *errmax,
                                           double *cfl_max, double *pe_max) {
  if (!S || !d_vol || !d_sav || !d_px || !d_py || !d_pz || !d_u || !d_v || !d_w || !d_pp ||
      !d_den || !d_b || (S->nf > 0 && (!d_coef || !d_scoef || !d_flux)))
    return TFS_ERR_INVALID;
  const int gc = grid_of(S->n), gf = grid_of(S->nf);
  double *part = S->part;
  hipLaunchKernelGGL(k_cv_cells, dim3(gc), dim3(kThreads), 0, nullptr, S->n, d_vol, d_sav, d_px,
                     d_py, d_pz, d_u, d_v, d_w);
  hipLaunchKernelGGL(k_cv_faces, dim3(gf), dim3(kThreads), 0, nullptr, S->n, S->nf, S->faces,
                     S->kind, d_pp, d_coef, d_bou, d_in_m, d_scoef, S->d, dt, visc, d_flux, d_u,
                     d_v, d_w, part + kParts, part + 2 * kParts);
  hipLaunchKernelGGL(k_cv_mass, dim3(gc), dim3(kThreads), 0, nullptr, S->n, S->ptr, S->inc,
                     d_flux, d_vol, d_den, d_b, part);
  hipLaunchKernelGGL(k_cv_final, dim3(1), dim3(kThreads), 0, nullptr, gf, gc, part,
                     part + 3 * kParts);
  TFS_G_HIP(hipGetLastError());
  double out[3];
  TFS_G_HIP(hipMemcpy(out, part + 3 * kParts, sizeof(out), hipMemcpyDeviceToHost));
  if (errmax) *errmax = out[0];
  if (cfl_max) *cfl_max = out[1];
  if (pe_max) *pe_max = out[2];
  return TFS_OK;
}

extern "C" int tfs_pressure_fractional_device(tfs_gradient *S, const double *d_f,
                                              const double *d_s, const double *d_vol,
                                              const double *d_sav, const double *d_u,
                                              const double *d_v, const double *d_w,
                                              const double *d_px, const double *d_py,
                                              const double *d_pz, const double *d_p,
                                              const double *d_den, double *d_flux,
                                              double *d_a12, double *d_b) {
  if (!S || S->has_buffer || !d_vol || !d_sav || !d_u || !d_v || !d_w || !d_px || !d_py ||
      !d_pz || !d_p || !d_den || !d_b ||
      (S->nf > 0 && (!d_f || !d_s || !d_flux || !d_a12)))
    return TFS_ERR_INVALID;
  const long long na = (long long)S->n + S->nb;
  // the per-cell records (48 bytes per cell slot) on the first call only:
  // meshes that never run this face loop do not hold them
  if (!S->rec && hipMalloc((void **)&S->rec, std::max<size_t>(6 * sizeof(double) * na, 8)) !=
                     hipSuccess) {
    (void)hipGetLastError();
    S->rec = nullptr;
    return TFS_ERR_NOMEM;
  }
  hipLaunchKernelGGL(k_pf_cells, dim3(grid_of(na)), dim3(kThreads), 0, nullptr, S->n, na, d_vol,
                     d_sav, d_u, d_v, d_w, d_px, d_py, d_pz, d_p, d_den, S->rec);
  if (S->nf > 0)
    hipLaunchKernelGGL(k_pf_faces, dim3(grid_of(S->nf)), dim3(kThreads), 0, nullptr, S->n, S->nf,
                       S->faces, d_f, d_s, S->rec, d_flux, d_a12);
  hipLaunchKernelGGL(k_pf_rhs, dim3(grid_of(S->n)), dim3(kThreads), 0, nullptr, S->n, S->ptr,
                     S->inc, d_flux, d_b);
  TFS_G_HIP(hipGetLastError());
  TFS_G_HIP(hipStreamSynchronize(nullptr));
  return TFS_OK;
}

extern "C" int tfs_pressure_simple_device(tfs_gradient *S, const double *d_f, const double *d_s,
                                          const double *d_vol, const double *d_sav,
                                          const double *d_u, const double *d_v,
                                          const double *d_w, const double *d_px,
                                          const double *d_py, const double *d_pz,
                                          const double *d_p, const double *d_den,
                                          double *d_flux, double *d_a12,
                                          unsigned char *d_bnd_on, double *d_b) {
  if (!S || S->has_buffer || !d_vol || !d_sav || !d_u || !d_v || !d_w || !d_px || !d_py ||
      !d_pz || !d_p || !d_den || !d_b ||
      (S->nf > 0 && (!d_f || !d_s || !d_flux || !d_a12 || !d_bnd_on)))
    return TFS_ERR_INVALID;
  if (S->nf > 0)
    hipLaunchKernelGGL(k_ps_faces, dim3(grid_of(S->nf)), dim3(kThreads), 0, nullptr, S->n, S->nf,
                       S->faces, S->kind, d_f, d_s, S->d, d_vol, d_sav, d_u, d_v, d_w, d_p, d_px,
                       d_py, d_pz, d_den, d_flux, d_a12, d_bnd_on);
  hipLaunchKernelGGL(k_pf_rhs, dim3(grid_of(S->n)), dim3(kThreads), 0, nullptr, S->n, S->ptr,
                     S->inc, d_flux, d_b);
  TFS_G_HIP(hipGetLastError());
  TFS_G_HIP(hipStreamSynchronize(nullptr));
  return TFS_OK;
}
