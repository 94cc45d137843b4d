// tfs_meshio.cpp -- T-FlowS mesh files (.cns connectivity, .geo geometry)
// and the fractional-step pressure matrix on a mesh read from them.
//
// The files are Fortran unformatted sequential files as the reference writes
// them (Sources/Shared/Save_Cns_Geo.f90) and reads them
// (Sources/Process/Load_Cns.f90:25-69, Sources/Shared/Load_Geo.f90:23-50):
// every record is a 4-byte length, the payload, and the length again;
// integers are 4 bytes, reals 8 (the reference is built with
// -fdefault-real-8).  The readers follow the record order of the reference's
// readers and check every record's length against the sizes read first.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "tfs_mesh.h"

namespace {

struct Records {
  FILE *f = nullptr;
  explicit Records(const char *path) : f(std::fopen(path, "rb")) {}
  ~Records() {
    if (f) std::fclose(f);
  }
  // Reads one record of exactly `bytes` bytes into dst (or skips it when dst
  // is null); false on a short file or a length that does not match.
  bool read(void *dst, size_t bytes) {
    uint32_t head = 0, tail = 0;
    if (std::fread(&head, 4, 1, f) != 1 || head != bytes) return false;
    if (dst) {
      if (bytes && std::fread(dst, 1, bytes, f) != bytes) return false;
    } else if (std::fseek(f, (long)bytes, SEEK_CUR) != 0) {
      return false;
    }
    return std::fread(&tail, 4, 1, f) == 1 && tail == head;
  }
  bool read_int(int *v) { return read(v, sizeof(int)); }
};

}  // namespace

extern "C" int tfs_mesh_read_cns(const char *path, int *n_cells, int *n_bnd_cells, int *n_faces,
                                 int *faces_c, int *material) {
  if (!path || !n_cells || !n_bnd_cells || !n_faces) return TFS_MESH_ERR_INVALID;
  Records r(path);
  if (!r.f) return TFS_MESH_ERR_IO;
  int nc = 0, nb = 0, nf = 0, n_sh = 0, n_mat = 0, n_bc = 0;
  // sizes (Load_Cns.f90:31-34), material and boundary-condition counts (:41-42)
  if (!r.read_int(&nc) || !r.read_int(&nb) || !r.read_int(&nf) || !r.read_int(&n_sh) ||
      !r.read_int(&n_mat) || !r.read_int(&n_bc) || nc < 1 || nb < 0 || nf < 0 || n_mat < 0 ||
      n_bc < 0)
    return TFS_MESH_ERR_FORMAT;
  *n_cells = nc;
  *n_bnd_cells = nb;
  *n_faces = nf;
  if (!faces_c && !material) return TFS_MESH_OK;
  // their names, 80 characters each (:48-53)
  for (int k = 0; k < n_mat + n_bc; ++k)
    if (!r.read(nullptr, 80)) return TFS_MESH_ERR_FORMAT;
  // cell materials, cells then boundary cells -1..-nb (:57-58)
  std::vector<int> mat((size_t)nc + nb);
  if (!r.read(mat.data(), sizeof(int) * (size_t)nc) ||
      !r.read(mat.data() + nc, sizeof(int) * (size_t)nb))
    return TFS_MESH_ERR_FORMAT;
  if (material) std::memcpy(material, mat.data(), sizeof(int) * mat.size());
  // faces: faces_c(1,:) then faces_c(2,:) (:61-62), stored here as (c1, c2)
  // pairs, i.e. the reference's faces_c(2, n_faces) in column-major order
  std::vector<int> c1((size_t)nf), c2((size_t)nf);
  if (!r.read(c1.data(), sizeof(int) * (size_t)nf) || !r.read(c2.data(), sizeof(int) * (size_t)nf))
    return TFS_MESH_ERR_FORMAT;
  for (int s = 0; s < nf; ++s) {
    if (c1[(size_t)s] < 1 || c1[(size_t)s] > nc || c2[(size_t)s] == 0 || c2[(size_t)s] > nc ||
        c2[(size_t)s] < -nb)
      return TFS_MESH_ERR_FORMAT;
    if (faces_c) {
      faces_c[2 * (size_t)s] = c1[(size_t)s];
      faces_c[2 * (size_t)s + 1] = c2[(size_t)s];
    }
  }
  return TFS_MESH_OK;
}

// Load_Cns.f90:62-69: after the faces, the boundary-condition mark of every
// boundary cell -1..-nb, then its copy cell (0 where it has none).
extern "C" int tfs_mesh_read_cns_marks(const char *path, int n_bnd_cells, int *bcmark,
                                       int *copy_c) {
  if (!path || n_bnd_cells < 0) return TFS_MESH_ERR_INVALID;
  int nc = 0, nb = 0, nf = 0;
  const int rc = tfs_mesh_read_cns(path, &nc, &nb, &nf, nullptr, nullptr);
  if (rc != TFS_MESH_OK) return rc;
  if (nb != n_bnd_cells) return TFS_MESH_ERR_INVALID;
  Records r(path);
  if (!r.f) return TFS_MESH_ERR_IO;
  int n_mat = 0, n_bc = 0, skip = 0;
  // the six sizes, the names, materials and faces, as tfs_mesh_read_cns reads them
  for (int k = 0; k < 4; ++k)
    if (!r.read_int(&skip)) return TFS_MESH_ERR_FORMAT;
  if (!r.read_int(&n_mat) || !r.read_int(&n_bc)) return TFS_MESH_ERR_FORMAT;
  for (int k = 0; k < n_mat + n_bc; ++k)
    if (!r.read(nullptr, 80)) return TFS_MESH_ERR_FORMAT;
  if (!r.read(nullptr, sizeof(int) * (size_t)nc) || !r.read(nullptr, sizeof(int) * (size_t)nb) ||
      !r.read(nullptr, sizeof(int) * (size_t)nf) || !r.read(nullptr, sizeof(int) * (size_t)nf))
    return TFS_MESH_ERR_FORMAT;
  std::vector<int> mark((size_t)nb), copy((size_t)nb);
  if (!r.read(mark.data(), sizeof(int) * (size_t)nb) ||
      !r.read(copy.data(), sizeof(int) * (size_t)nb))
    return TFS_MESH_ERR_FORMAT;
  if (bcmark) std::memcpy(bcmark, mark.data(), sizeof(int) * mark.size());
  if (copy_c) std::memcpy(copy_c, copy.data(), sizeof(int) * copy.size());
  return TFS_MESH_OK;
}

extern "C" int tfs_mesh_read_geo(const char *path, int n_cells, int n_bnd_cells, int n_faces,
                                 double *xc, double *vol, double *s, double *shift, double *f,
                                 double *xf) {
  if (!path || n_cells < 1 || n_bnd_cells < 0 || n_faces < 0) return TFS_MESH_ERR_INVALID;
  Records r(path);
  if (!r.f) return TFS_MESH_ERR_IO;
  const size_t nc = (size_t)n_cells, nb = (size_t)n_bnd_cells, nf = (size_t)n_faces;
  const size_t na = nc + nb;
  // each record to its slot in `dst` (or skipped when dst is null)
  auto rec = [&](double *dst, size_t count) { return r.read(dst, sizeof(double) * count); };
  // cell centres: x, y, z of the cells, then of the boundary cells (:25-31)
  for (int k = 0; k < 3; ++k)
    if (!rec(xc ? xc + k * na : nullptr, nc)) return TFS_MESH_ERR_FORMAT;
  for (int k = 0; k < 3; ++k)
    if (!rec(xc ? xc + k * na + nc : nullptr, nb)) return TFS_MESH_ERR_FORMAT;
  // volumes, then delta and the wall distance (not kept) (:33-35)
  if (!rec(vol, nc) || !rec(nullptr, nc) || !rec(nullptr, nc)) return TFS_MESH_ERR_FORMAT;
  // face surfaces, periodic shifts, weights, face centres (:37-48)
  for (int k = 0; k < 3; ++k)
    if (!rec(s ? s + k * nf : nullptr, nf)) return TFS_MESH_ERR_FORMAT;
  for (int k = 0; k < 3; ++k)
    if (!rec(shift ? shift + k * nf : nullptr, nf)) return TFS_MESH_ERR_FORMAT;
  if (!rec(f, nf)) return TFS_MESH_ERR_FORMAT;
  for (int k = 0; k < 3; ++k)
    if (!rec(xf ? xf + k * nf : nullptr, nf)) return TFS_MESH_ERR_FORMAT;
  return TFS_MESH_OK;
}

// Compute_Geometry.f90:22-49: Scoef = |S|^2 / (d . S), with d the distance
// from the centre of c1 to the centre of c2 moved by the face's periodic
// shift; the operations in the reference's order.
extern "C" int tfs_mesh_face_coef(int n_cells, int n_bnd_cells, int n_faces, const int *faces_c,
                                  const double *xc, const double *s, const double *shift,
                                  double *d, double *scoef) {
  if (n_cells < 1 || n_bnd_cells < 0 || n_faces < 0 ||
      (n_faces > 0 && (!faces_c || !xc || !s || !shift || !scoef)))
    return TFS_MESH_ERR_INVALID;
  const size_t na = (size_t)n_cells + (size_t)n_bnd_cells, nf = (size_t)n_faces;
  // position in xc of cell c (1..n_cells) or boundary cell c (-1..-n_bnd)
  auto at = [&](int c) -> size_t {
    return c > 0 ? (size_t)(c - 1) : (size_t)n_cells + (size_t)(-c - 1);
  };
  for (size_t f = 0; f < nf; ++f) {
    const int c1 = faces_c[2 * f], c2 = faces_c[2 * f + 1];
    if (c1 < 1 || c1 > n_cells || c2 == 0 || c2 > n_cells || c2 < -n_bnd_cells)
      return TFS_MESH_ERR_INVALID;
    const double sx = s[f], sy = s[nf + f], sz = s[2 * nf + f];
    const double area2 = sx * sx + sy * sy + sz * sz;
    const size_t i1 = at(c1), i2 = at(c2);
    const double dx = (xc[i2] + shift[f]) - xc[i1];
    const double dy = (xc[na + i2] + shift[nf + f]) - xc[na + i1];
    const double dz = (xc[2 * na + i2] + shift[2 * nf + f]) - xc[2 * na + i1];
    if (d) {
      d[f] = dx;
      d[nf + f] = dy;
      d[2 * nf + f] = dz;
    }
    scoef[f] = area2 / (dx * sx + dy * sy + dz * sz);
  }
  return TFS_MESH_OK;
}

// Pressure_Matrix_Fractional.f90:19-48: val := 0; for every interior face the
// two off-diagonals are -dt Scoef and both diagonals gain dt Scoef, in face
// order; a boundary face adds dt Scoef to its cell's diagonal only where
// bnd_on[s] is set (the reference does so for BUFFER faces of a decomposed
// run).  pos / dia as tfs_topology_build writes them, with index base `base`.
extern "C" int tfs_mesh_pressure_values(int n_cells, int64_t n_faces, const int *faces_c,
                                        const int *pos, const int *dia, int base, int64_t nnz,
                                        const double *scoef, double dt,
                                        const unsigned char *bnd_on, double *val) {
  if (n_cells < 1 || n_faces < 0 || nnz < n_cells || !dia || !val || (base != 0 && base != 1) ||
      (n_faces > 0 && (!faces_c || !pos || !scoef)))
    return TFS_MESH_ERR_INVALID;
  for (int64_t k = 0; k < nnz; ++k) val[k] = 0.0;
  auto ok = [&](int64_t q) { return q >= 0 && q < nnz; };
  for (int64_t f = 0; f < n_faces; ++f) {
    const int c1 = faces_c[2 * f], c2 = faces_c[2 * f + 1];
    if (c1 < 1 || c1 > n_cells || c2 > n_cells) return TFS_MESH_ERR_INVALID;
    const double a12 = dt * scoef[f];
    const int64_t d1 = (int64_t)dia[c1 - 1] - base;
    if (!ok(d1)) return TFS_MESH_ERR_INVALID;
    if (c2 > 0) {
      const int64_t p1 = (int64_t)pos[2 * f] - base, p2 = (int64_t)pos[2 * f + 1] - base;
      const int64_t d2 = (int64_t)dia[c2 - 1] - base;
      if (!ok(p1) || !ok(p2) || !ok(d2)) return TFS_MESH_ERR_INVALID;
      val[p1] = -a12;
      val[p2] = -a12;
      val[d1] = val[d1] + a12;
      val[d2] = val[d2] + a12;
    } else if (bnd_on && bnd_on[f]) {
      val[d1] = val[d1] + a12;
    }
  }
  return TFS_MESH_OK;
}
