// tfs_mesh.cpp -- CRS topology from faces and synthetic hex-mesh cases.
//
// Host code (g++), built into libtfs_mesh.so.  The topology builder follows
// Sources/Process/Matrix_Mod/Topology.f90:33-121: one stencil slot per cell
// plus one per interior face side, diagonal placed first, columns sorted
// ascending within each row, dia = position of the diagonal, pos(1:2,s) =
// positions of (c1,c2) and (c2,c1).  Because the columns end up sorted, the
// layout does not depend on face order.
#include "tfs_mesh.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Stream key for (seed, tag); draw i is mix64(key + (i+1)*golden), which is
// exactly the i-th output of a SplitMix64 generator started at `key`.
inline uint64_t stream_key(uint64_t seed, uint64_t tag) {
  return mix64(seed ^ (tag * 0xD1B54A32D192ED03ull));
}
inline double u01_at(uint64_t key, uint64_t i) {
  return (double)(mix64(key + (i + 1) * kGolden) >> 11) * 0x1.0p-53;
}

enum : uint64_t { TAG_RHS = 1, TAG_PERM = 2, TAG_JITTER = 3, TAG_VEL = 4 };

struct V3 {
  double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// 1-D tanh clustering towards both ends:
// x_i = (1 + tanh(beta (2i/n - 1)) / tanh(beta)) / 2, the symmetric case of
// Sources/Utilities/Stretch.f90:57-63 with pr = tanh(beta).
std::vector<double> stretch(int n, double beta) {
  std::vector<double> x(n + 1);
  for (int i = 0; i <= n; ++i) {
    const double xi = 2.0 * i / n - 1.0;
    x[i] = beta > 1e-12 ? 0.5 * (1.0 + std::tanh(beta * xi) / std::tanh(beta))
                        : 0.5 * (1.0 + xi);
  }
  x[0] = 0.0;
  x[n] = 1.0;
  return x;
}

// Divergence-free velocity from a vector potential psi = sum_m a_m sin(2 pi k_m.x + phi_m);
// the face flux is the circulation of psi around the face (Stokes), so the
// outward fluxes of every closed cell cancel edge by edge.
struct Potential {
  int n = 0;
  double k[16][3], a[16][3], phi[16];
  V3 eval(V3 p) const {
    V3 s{0, 0, 0};
    for (int m = 0; m < n; ++m) {
      const double arg = 2.0 * M_PI * (k[m][0] * p.x + k[m][1] * p.y + k[m][2] * p.z) + phi[m];
      const double sn = std::sin(arg);
      s = s + sn * V3{a[m][0], a[m][1], a[m][2]};
    }
    return s;
  }
};

}  // namespace

struct tfs_case {
  tfs_case_params p;
  int n = 0;
  int64_t nf = 0;
  std::vector<int> faces;     // 2*nf, 1-based c1, c2 (negative: boundary)
  std::vector<double> a12;    // per face: coefficient of c2 in row c1 (or boundary diag add)
  std::vector<double> a21;    // per face: coefficient of c1 in row c2
  std::vector<double> dextra; // per cell: inertia V/dt
  std::vector<double> b;
  double dt = 0.0;
  double max_aspect = 1.0;
};

extern "C" double tfs_rng_u01(uint64_t seed, uint64_t tag, uint64_t i) {
  return u01_at(stream_key(seed, tag), i);
}

extern "C" int64_t tfs_topology_nnz(int n_cells, int64_t n_faces, const int *faces_c) {
  int64_t n = n_cells;
  for (int64_t s = 0; s < n_faces; ++s)
    if (faces_c[2 * s + 1] > 0) n += 2;
  return n;
}

extern "C" int tfs_topology_build(int n_cells, int64_t n_faces, const int *faces_c,
                                  int base, int *row, int *col, int *dia, int *pos) {
  if (n_cells < 0 || (base != 0 && base != 1)) return -1;
  const int64_t nnz = tfs_topology_nnz(n_cells, n_faces, faces_c);
  if (nnz > INT_MAX - 1) return -2;
  std::vector<int> width(n_cells, 1);  // Topology.f90:26,34-41
  for (int64_t s = 0; s < n_faces; ++s) {
    const int c1 = faces_c[2 * s], c2 = faces_c[2 * s + 1];
    if (c1 < 1 || c1 > n_cells || c2 > n_cells) return -3;
    if (c2 > 0) {
      width[c1 - 1]++;
      width[c2 - 1]++;
    }
  }
  // row starts (0-based here), diagonal first (Topology.f90:54-59)
  std::vector<int> fill(n_cells, 1);
  row[0] = 0;
  for (int c = 0; c < n_cells; ++c) {
    row[c + 1] = row[c] + width[c];
    col[row[c]] = c;
  }
  // off-diagonals in face order (Topology.f90:62-71)
  for (int64_t s = 0; s < n_faces; ++s) {
    const int c1 = faces_c[2 * s] - 1, c2 = faces_c[2 * s + 1] - 1;
    if (c2 >= 0) {
      col[row[c1] + fill[c1]++] = c2;
      col[row[c2] + fill[c2]++] = c1;
    }
  }
  // ascending columns per row, diagonal position (Topology.f90:75-82)
  for (int c = 0; c < n_cells; ++c) {
    int *b = col + row[c], *e = col + row[c + 1];
    if (e - b <= 16) {
      for (int *q = b + 1; q < e; ++q)  // insertion sort, short rows
        for (int *r = q; r > b && r[-1] > r[0]; --r) std::swap(r[-1], r[0]);
    } else {
      std::sort(b, e);
    }
    for (int q = row[c]; q < row[c + 1]; ++q)
      if (col[q] == c) dia[c] = q;
  }
  // face -> entry positions (Topology.f90:106-121)
  if (pos) {
    for (int64_t s = 0; s < n_faces; ++s) {
      const int c1 = faces_c[2 * s] - 1, c2 = faces_c[2 * s + 1] - 1;
      pos[2 * s] = pos[2 * s + 1] = -1;
      if (c2 >= 0) {
        for (int q = row[c1]; q < row[c1 + 1]; ++q)
          if (col[q] == c2) pos[2 * s] = q;
        for (int q = row[c2]; q < row[c2 + 1]; ++q)
          if (col[q] == c1) pos[2 * s + 1] = q;
      }
    }
  }
  if (base == 1) {
    for (int c = 0; c <= n_cells; ++c) row[c] += 1;
    for (int64_t q = 0; q < nnz; ++q) col[q] += 1;
    for (int c = 0; c < n_cells; ++c) dia[c] += 1;
    if (pos)
      for (int64_t s = 0; s < 2 * n_faces; ++s) pos[s] += 1;
  }
  return 0;
}

extern "C" int tfs_case_create(const tfs_case_params *pp, tfs_case **out) {
  *out = nullptr;
  const tfs_case_params &p = *pp;
  if (p.nx < 1 || p.ny < 1 || p.nz < 1) return -1;
  if (p.n_modes < 0 || p.n_modes > 16) return -1;
  const int64_t ncell = (int64_t)p.nx * p.ny * p.nz;
  if (ncell > INT_MAX / 8) return -2;
  tfs_case *c = new (std::nothrow) tfs_case;
  if (!c) return -3;
  c->p = p;
  c->n = (int)ncell;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const int n = c->n;

  // Node coordinates: stretched lines plus jitter of interior nodes.
  const std::vector<double> X = stretch(nx, p.beta[0]), Y = stretch(ny, p.beta[1]),
                            Z = stretch(nz, p.beta[2]);
  auto hloc = [](const std::vector<double> &g, int i) {
    return std::min(g[i] - g[i - 1], g[i + 1] - g[i]);
  };
  const int64_t NX = nx + 1, NY = ny + 1, NZ = nz + 1;
  std::vector<V3> node((size_t)(NX * NY * NZ));
  const uint64_t kj = stream_key(p.seed, TAG_JITTER);
  for (int64_t k = 0; k < NZ; ++k)
    for (int64_t j = 0; j < NY; ++j)
      for (int64_t i = 0; i < NX; ++i) {
        const int64_t id = i + NX * (j + NY * k);
        V3 q{X[i], Y[j], Z[k]};
        if (p.jitter > 0) {
          if (i > 0 && i < nx) q.x += p.jitter * hloc(X, (int)i) * (2 * u01_at(kj, 3 * id) - 1);
          if (j > 0 && j < ny) q.y += p.jitter * hloc(Y, (int)j) * (2 * u01_at(kj, 3 * id + 1) - 1);
          if (k > 0 && k < nz) q.z += p.jitter * hloc(Z, (int)k) * (2 * u01_at(kj, 3 * id + 2) - 1);
        }
        node[(size_t)id] = q;
      }
  auto N_ = [&](int64_t i, int64_t j, int64_t k) -> const V3 & {
    return node[(size_t)(i + NX * (j + NY * k))];
  };

  // Cell numbering: natural (i fastest) or a seeded Fisher-Yates shuffle.
  std::vector<int> id(n);
  for (int q = 0; q < n; ++q) id[q] = q;
  if (p.permute) {
    const uint64_t kp = stream_key(p.seed, TAG_PERM);
    for (int q = n - 1; q > 0; --q) {
      int r = (int)(u01_at(kp, (uint64_t)(n - 1 - q)) * (q + 1));
      if (r > q) r = q;
      std::swap(id[q], id[r]);
    }
  }
  auto cell = [&](int64_t i, int64_t j, int64_t k) { return id[(size_t)(i + nx * (j + (int64_t)ny * k))]; };

  // The four corners of a face, in the circulation order whose right-hand
  // normal points along +axis; S = (c - a) x (d - b) / 2.
  auto corners = [&](int dir, int64_t i, int64_t j, int64_t k, V3 q[4]) {
    if (dir == 0) { q[0] = N_(i, j, k); q[1] = N_(i, j + 1, k); q[2] = N_(i, j + 1, k + 1); q[3] = N_(i, j, k + 1); }
    if (dir == 1) { q[0] = N_(i, j, k); q[1] = N_(i, j, k + 1); q[2] = N_(i + 1, j, k + 1); q[3] = N_(i + 1, j, k); }
    if (dir == 2) { q[0] = N_(i, j, k); q[1] = N_(i + 1, j, k); q[2] = N_(i + 1, j + 1, k); q[3] = N_(i, j + 1, k); }
  };
  auto area = [](const V3 q[4]) { return 0.5 * cross(q[2] - q[0], q[3] - q[1]); };
  auto centre4 = [](const V3 q[4]) { return 0.25 * (q[0] + q[1] + q[2] + q[3]); };

  // Cell centres (mean of corners), volumes (divergence theorem), aspect.
  std::vector<V3> xc((size_t)n);
  std::vector<double> vol((size_t)n);
  double amax = 1.0;
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        V3 s{0, 0, 0};
        for (int t = 0; t < 8; ++t) s = s + N_(i + (t & 1), j + ((t >> 1) & 1), k + (t >> 2));
        const int c = cell(i, j, k);
        xc[(size_t)c] = 0.125 * s;
        double v = 0.0;
        V3 q[4];
        for (int dir = 0; dir < 3; ++dir) {
          const int64_t di = dir == 0, dj = dir == 1, dk = dir == 2;
          corners(dir, i + di, j + dj, k + dk, q);  // + face, outward along +axis
          v += dot(area(q), centre4(q));
          corners(dir, i, j, k, q);                 // - face, outward along -axis
          v -= dot(area(q), centre4(q));
        }
        vol[(size_t)c] = v / 3.0;
        double emin = 1e300, emax = 0.0;
        for (int t = 0; t < 8; ++t) {
          const int64_t ti = t & 1, tj = (t >> 1) & 1, tk = t >> 2;
          const V3 a = N_(i + ti, j + tj, k + tk);
          if (!ti) { const double e = std::sqrt(dot(N_(i + 1, j + tj, k + tk) - a, N_(i + 1, j + tj, k + tk) - a)); emin = std::min(emin, e); emax = std::max(emax, e); }
          if (!tj) { const double e = std::sqrt(dot(N_(i + ti, j + 1, k + tk) - a, N_(i + ti, j + 1, k + tk) - a)); emin = std::min(emin, e); emax = std::max(emax, e); }
          if (!tk) { const double e = std::sqrt(dot(N_(i + ti, j + tj, k + 1) - a, N_(i + ti, j + tj, k + 1) - a)); emin = std::min(emin, e); emax = std::max(emax, e); }
        }
        amax = std::max(amax, emax / emin);
      }
  c->max_aspect = amax;

  const bool mom = p.kind == TFS_CASE_MOMENTUM;
  Potential pot;
  if (mom) {
    const uint64_t kv = stream_key(p.seed, TAG_VEL);
    uint64_t t = 0;
    pot.n = p.n_modes;
    double bound = 0.0;
    for (int m = 0; m < pot.n; ++m) {
      do {
        for (int d = 0; d < 3; ++d) pot.k[m][d] = std::floor(u01_at(kv, t++) * 5.0) - 2.0;
      } while (pot.k[m][0] == 0 && pot.k[m][1] == 0 && pot.k[m][2] == 0);
      for (int d = 0; d < 3; ++d) pot.a[m][d] = 2 * u01_at(kv, t++) - 1;
      pot.phi[m] = 2 * M_PI * u01_at(kv, t++);
      const V3 km{pot.k[m][0], pot.k[m][1], pot.k[m][2]}, am{pot.a[m][0], pot.a[m][1], pot.a[m][2]};
      const V3 ka = cross(km, am);
      bound += 2 * M_PI * std::sqrt(dot(ka, ka));
    }
    // |u| = |curl psi| <= sum_m 2 pi |k_m x a_m|; scale so that |u| <= 1.
    for (int m = 0; m < pot.n; ++m)
      for (int d = 0; d < 3; ++d) pot.a[m][d] /= bound > 0 ? bound : 1.0;
    double hmin = 1e300;
    for (int i = 0; i < nx; ++i) hmin = std::min(hmin, X[i + 1] - X[i]);
    for (int j = 0; j < ny; ++j) hmin = std::min(hmin, Y[j + 1] - Y[j]);
    for (int k = 0; k < nz; ++k) hmin = std::min(hmin, Z[k + 1] - Z[k]);
    hmin *= (1.0 - 2.0 * p.jitter);
    c->dt = p.cfl * hmin / 1.0;
  }
  auto flux = [&](const V3 q[4]) {  // circulation of psi: sum over edges psi(mid).(q_b - q_a)
    double f = 0.0;
    for (int e = 0; e < 4; ++e) {
      const V3 a = q[e], b = q[(e + 1) & 3];
      f += dot(pot.eval(0.5 * (a + b)), b - a);
    }
    return f;
  };

  // Faces: x-, y-, then z-oriented, node planes along the axis outermost.
  c->nf = (int64_t)(nx + 1) * ny * nz + (int64_t)nx * (ny + 1) * nz + (int64_t)nx * ny * (nz + 1);
  c->faces.resize((size_t)(2 * c->nf));
  c->a12.resize((size_t)c->nf);
  c->a21.resize((size_t)c->nf);
  int64_t s = 0;
  int nbnd = 0;
  const int ncell_axis[3] = {nx, ny, nz};
  for (int dir = 0; dir < 3; ++dir) {
    const int64_t ei = nx + (dir == 0), ej = ny + (dir == 1), ek = nz + (dir == 2);
    for (int64_t k = 0; k < ek; ++k)
      for (int64_t j = 0; j < ej; ++j)
        for (int64_t i = 0; i < ei; ++i, ++s) {
          const int64_t plane = dir == 0 ? i : dir == 1 ? j : k;
          const int64_t di = dir == 0, dj = dir == 1, dk = dir == 2;
          V3 q[4];
          corners(dir, i, j, k, q);
          V3 S = area(q);
          const V3 fc = centre4(q);
          double F = mom ? flux(q) : 0.0;
          int c1, c2;
          V3 d;
          if (plane == 0 || plane == ncell_axis[dir]) {
            // boundary face: c1 is the interior cell, S and F point outward
            c1 = plane == 0 ? cell(i, j, k) : cell(i - di, j - dj, k - dk);
            if (plane == 0) { S = -1.0 * S; F = -F; }
            c2 = -(++nbnd);
            d = fc - xc[(size_t)c1];
          } else {
            c1 = cell(i - di, j - dj, k - dk);
            c2 = cell(i, j, k);
            d = xc[(size_t)c2] - xc[(size_t)c1];
          }
          // Scoef = |S|^2 / (S.d), Compute_Geometry.f90:22-49
          const double scoef = dot(S, S) / dot(S, d);
          c->faces[(size_t)(2 * s)] = c1 + 1;
          c->faces[(size_t)(2 * s + 1)] = c2 >= 0 ? c2 + 1 : c2;
          if (!mom) {
            // Pressure_Matrix_Fractional.f90:31-38 (dt = 1); walls add nothing
            c->a12[(size_t)s] = c2 >= 0 ? scoef : 0.0;
            c->a21[(size_t)s] = c2 >= 0 ? scoef : 0.0;
          } else {
            // NewUVW.f90:349-384, fully implicit diffusion + upwind
            const double A0 = p.nu * scoef;
            if (c2 >= 0) {
              c->a12[(size_t)s] = A0 - std::min(F, 0.0);
              c->a21[(size_t)s] = A0 + std::max(F, 0.0);
            } else {
              c->a12[(size_t)s] = A0;  // no-slip wall: zero flux, diffusion only
              c->a21[(size_t)s] = 0.0;
            }
          }
        }
  }
  c->dextra.assign((size_t)n, 0.0);
  if (mom)  // inertia, NewUVW.f90:500-506 (rho = 1)
    for (int q = 0; q < n; ++q) c->dextra[(size_t)q] = vol[(size_t)q] / c->dt;

  const uint64_t kr = stream_key(p.seed, TAG_RHS);
  c->b.resize((size_t)n);
  double mean = 0.0;
  for (int q = 0; q < n; ++q) {
    c->b[(size_t)q] = 2 * u01_at(kr, (uint64_t)q) - 1;
    mean += c->b[(size_t)q];
  }
  if (p.rhs_zero_mean) {
    mean /= n;
    for (int q = 0; q < n; ++q) c->b[(size_t)q] -= mean;
  }
  *out = c;
  return 0;
}

extern "C" void tfs_case_destroy(tfs_case *c) { delete c; }

extern "C" void tfs_case_sizes(const tfs_case *c, int *n_cells, int64_t *n_faces, int64_t *nnz) {
  if (n_cells) *n_cells = c->n;
  if (n_faces) *n_faces = c->nf;
  if (nnz) *nnz = tfs_topology_nnz(c->n, c->nf, c->faces.data());
}

extern "C" int tfs_case_faces(const tfs_case *c, int *faces_c) {
  std::memcpy(faces_c, c->faces.data(), sizeof(int) * c->faces.size());
  return 0;
}

extern "C" int tfs_case_face_coefs(const tfs_case *c, double *a12, double *a21, double *dextra) {
  std::memcpy(a12, c->a12.data(), sizeof(double) * c->a12.size());
  std::memcpy(a21, c->a21.data(), sizeof(double) * c->a21.size());
  std::memcpy(dextra, c->dextra.data(), sizeof(double) * c->dextra.size());
  return 0;
}

extern "C" double tfs_case_dt(const tfs_case *c) { return c->dt; }
extern "C" double tfs_case_max_aspect(const tfs_case *c) { return c->max_aspect; }

extern "C" int tfs_case_crs(const tfs_case *c, int base, int *row, int *col, int *dia,
                            double *val, double *b) {
  const int n = c->n;
  std::vector<int> pos((size_t)(2 * c->nf));
  int rc = tfs_topology_build(n, c->nf, c->faces.data(), 0, row, col, dia, pos.data());
  if (rc) return rc;
  const int64_t nnz = row[n];
  std::fill(val, val + nnz, 0.0);
  // Face loop scatter through pos/dia, as the reference assembly does.
  for (int64_t s = 0; s < c->nf; ++s) {
    const int c1 = c->faces[(size_t)(2 * s)] - 1, c2 = c->faces[(size_t)(2 * s + 1)];
    if (c2 > 0) {
      val[pos[(size_t)(2 * s)]] -= c->a12[(size_t)s];
      val[dia[c1]] += c->a12[(size_t)s];
      val[pos[(size_t)(2 * s + 1)]] -= c->a21[(size_t)s];
      val[dia[c2 - 1]] += c->a21[(size_t)s];
    } else {
      val[dia[c1]] += c->a12[(size_t)s];
    }
  }
  for (int q = 0; q < n; ++q) val[dia[q]] += c->dextra[(size_t)q];
  if (c->p.kind == TFS_CASE_POISSON && c->p.pin_first && n > 0) val[dia[0]] *= 2.0;
  std::memcpy(b, c->b.data(), sizeof(double) * (size_t)n);
  if (base == 1) {
    for (int q = 0; q <= n; ++q) row[q] += 1;
    for (int64_t q = 0; q < nnz; ++q) col[q] += 1;
    for (int q = 0; q < n; ++q) dia[q] += 1;
  }
  return 0;
}
