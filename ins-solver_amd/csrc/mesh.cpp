#include "mesh.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <random>
#include <stdexcept>
#include <string>

#include "refelem.h"

namespace insp {

namespace {

inline double uniform01(std::mt19937_64& g) { return double(g() >> 11) * 0x1.0p-53; }

double signed_measure(const HostMesh& m, int64_t e) {
  const int d = m.dim;
  const int* c = &m.cell2nodes[(size_t)e * (d + 1)];
  if (d == 2) {
    const double* a = &m.coords[(size_t)c[0] * 2];
    const double* b = &m.coords[(size_t)c[1] * 2];
    const double* q = &m.coords[(size_t)c[2] * 2];
    return (b[0] - a[0]) * (q[1] - a[1]) - (q[0] - a[0]) * (b[1] - a[1]);
  }
  const double* a = &m.coords[(size_t)c[0] * 3];
  double u[3][3];
  for (int k = 0; k < 3; ++k)
    for (int x = 0; x < 3; ++x) u[k][x] = m.coords[(size_t)c[k + 1] * 3 + x] - a[x];
  return u[0][0] * (u[1][1] * u[2][2] - u[1][2] * u[2][1]) -
         u[0][1] * (u[1][0] * u[2][2] - u[1][2] * u[2][0]) +
         u[0][2] * (u[1][0] * u[2][1] - u[1][1] * u[2][0]);
}

}  // namespace

HostMesh make_square_mesh(int nx, int ny, double jitter, uint64_t seed) {
  if (nx < 1 || ny < 1) throw std::runtime_error("square mesh needs nx, ny >= 1");
  HostMesh m;
  m.dim = 2;
  m.n_nodes = (int64_t)(nx + 1) * (ny + 1);
  m.coords.resize((size_t)m.n_nodes * 2);
  std::mt19937_64 g(seed);
  const double hx = 1.0 / nx, hy = 1.0 / ny;
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      const size_t v = (size_t)j * (nx + 1) + i;
      double x = i * hx, y = j * hy;
      if (i > 0 && i < nx && j > 0 && j < ny) {
        x += (2 * uniform01(g) - 1) * jitter * hx;
        y += (2 * uniform01(g) - 1) * jitter * hy;
      }
      m.coords[v * 2] = x;
      m.coords[v * 2 + 1] = y;
    }
  m.n_cells = (int64_t)2 * nx * ny;
  m.cell2nodes.resize((size_t)m.n_cells * 3);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int v00 = j * (nx + 1) + i, v10 = v00 + 1, v01 = v00 + nx + 1, v11 = v01 + 1;
      const size_t e = 2 * ((size_t)j * nx + i);
      const int t0[3] = {v00, v10, v11}, t1[3] = {v00, v11, v01};
      std::copy(t0, t0 + 3, &m.cell2nodes[e * 3]);
      std::copy(t1, t1 + 3, &m.cell2nodes[(e + 1) * 3]);
    }
  build_topology(m);
  build_geometry(m);
  return m;
}

HostMesh make_cube_mesh(int n, double jitter, uint64_t seed) {
  if (n < 1) throw std::runtime_error("cube mesh needs n >= 1");
  HostMesh m;
  m.dim = 3;
  const int n1 = n + 1;
  m.n_nodes = (int64_t)n1 * n1 * n1;
  m.coords.resize((size_t)m.n_nodes * 3);
  std::mt19937_64 g(seed);
  const double h = 1.0 / n;
  for (int k = 0; k <= n; ++k)
    for (int j = 0; j <= n; ++j)
      for (int i = 0; i <= n; ++i) {
        const size_t v = ((size_t)k * n1 + j) * n1 + i;
        double x[3] = {i * h, j * h, k * h};
        if (i > 0 && i < n && j > 0 && j < n && k > 0 && k < n)
          for (int c = 0; c < 3; ++c) x[c] += (2 * uniform01(g) - 1) * jitter * h;
        for (int c = 0; c < 3; ++c) m.coords[v * 3 + c] = x[c];
      }
  m.n_cells = (int64_t)6 * n * n * n;
  m.cell2nodes.resize((size_t)m.n_cells * 4);
  static const int axis_perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2},
                                      {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const size_t cube = ((size_t)k * n + j) * n + i;
        for (int t = 0; t < 6; ++t) {
          int p[3] = {i, j, k};
          int verts[4];
          verts[0] = (p[2] * n1 + p[1]) * n1 + p[0];
          for (int s = 0; s < 3; ++s) {
            p[axis_perm[t][s]] += 1;
            verts[s + 1] = (p[2] * n1 + p[1]) * n1 + p[0];
          }
          std::copy(verts, verts + 4, &m.cell2nodes[(cube * 6 + t) * 4]);
        }
      }
  build_topology(m);
  build_geometry(m);
  return m;
}

void build_topology(HostMesh& m) {
  const int d = m.dim, nv = d + 1, Nf = d + 1;
  const int* fv = face_vertices(d);
  // Orient every cell positively.
  for (int64_t e = 0; e < m.n_cells; ++e) {
    const double vol = signed_measure(m, e);
    if (vol == 0.0) throw std::runtime_error("degenerate cell");
    if (vol < 0) std::swap(m.cell2nodes[(size_t)e * nv + d - 1], m.cell2nodes[(size_t)e * nv + d]);
  }
  struct FaceRec {
    std::array<int, 3> key;
    int64_t cell;
    int face;
  };
  std::vector<FaceRec> recs((size_t)m.n_cells * Nf);
  for (int64_t e = 0; e < m.n_cells; ++e)
    for (int f = 0; f < Nf; ++f) {
      FaceRec& r = recs[(size_t)e * Nf + f];
      r.key = {-1, -1, -1};
      for (int k = 0; k < d; ++k) r.key[k] = m.cell2nodes[(size_t)e * nv + fv[f * d + k]];
      std::sort(r.key.begin(), r.key.begin() + d);
      r.cell = e;
      r.face = f;
    }
  std::sort(recs.begin(), recs.end(), [](const FaceRec& a, const FaceRec& b) {
    if (a.key != b.key) return a.key < b.key;
    if (a.cell != b.cell) return a.cell < b.cell;
    return a.face < b.face;
  });
  m.nbr.assign((size_t)m.n_cells * Nf, INT32_MIN);
  m.nbr_face.assign((size_t)m.n_cells * Nf, 0);
  m.nbr_perm.assign((size_t)m.n_cells * Nf, 0);
  const int nperm = d == 2 ? 2 : 6;
  static const int perms2[2][3] = {{0, 1, 0}, {1, 0, 0}};
  static const int perms3[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  auto find_perm = [&](int64_t e, int f, int64_t en, int fn) {
    int mine[3], theirs[3];
    for (int k = 0; k < d; ++k) {
      mine[k] = m.cell2nodes[(size_t)e * nv + fv[f * d + k]];
      theirs[k] = m.cell2nodes[(size_t)en * nv + fv[fn * d + k]];
    }
    for (int p = 0; p < nperm; ++p) {
      bool ok = true;
      for (int k = 0; k < d; ++k) {
        const int src = d == 2 ? perms2[p][k] : perms3[p][k];
        ok &= theirs[k] == mine[src];
      }
      if (ok) return p;
    }
    throw std::runtime_error("face vertex permutation not found");
  };
  std::vector<std::array<int64_t, 4>> interior;  // e1, f1, e2, f2
  std::vector<std::array<int64_t, 2>> boundary;
  size_t i = 0;
  while (i < recs.size()) {
    size_t j = i + 1;
    while (j < recs.size() && recs[j].key == recs[i].key) ++j;
    if (j - i == 1) {
      boundary.push_back({recs[i].cell, recs[i].face});
    } else if (j - i == 2) {
      interior.push_back({recs[i].cell, recs[i].face, recs[i + 1].cell, recs[i + 1].face});
    } else {
      throw std::runtime_error("non-manifold face shared by more than two cells");
    }
    i = j;
  }
  std::sort(interior.begin(), interior.end());
  std::sort(boundary.begin(), boundary.end());
  m.n_faces = (int64_t)interior.size();
  m.face2cells.resize((size_t)m.n_faces * 2);
  m.faceNum.resize((size_t)m.n_faces * 2);
  for (size_t k = 0; k < interior.size(); ++k) {
    const int64_t e1 = interior[k][0], e2 = interior[k][2];
    const int f1 = (int)interior[k][1], f2 = (int)interior[k][3];
    m.face2cells[k * 2] = (int)e1;
    m.face2cells[k * 2 + 1] = (int)e2;
    m.faceNum[k * 2] = f1;
    m.faceNum[k * 2 + 1] = f2;
    m.nbr[(size_t)e1 * Nf + f1] = (int)e2;
    m.nbr_face[(size_t)e1 * Nf + f1] = f2;
    m.nbr_perm[(size_t)e1 * Nf + f1] = find_perm(e1, f1, e2, f2);
    m.nbr[(size_t)e2 * Nf + f2] = (int)e1;
    m.nbr_face[(size_t)e2 * Nf + f2] = f1;
    m.nbr_perm[(size_t)e2 * Nf + f2] = find_perm(e2, f2, e1, f1);
  }
  m.n_bfaces = (int64_t)boundary.size();
  m.bface2cells.resize(m.n_bfaces);
  m.bfaceNum.resize(m.n_bfaces);
  for (size_t k = 0; k < boundary.size(); ++k) {
    m.bface2cells[k] = (int)boundary[k][0];
    m.bfaceNum[k] = (int)boundary[k][1];
    m.nbr[(size_t)boundary[k][0] * Nf + boundary[k][1]] = -(int)k - 1;
  }
}

int64_t first_negative_cell(const HostMesh& m) {
  for (int64_t e = 0; e < m.n_cells; ++e)
    if (signed_measure(m, e) < 0) return e;
  return -1;
}

namespace {
// Permutation p (the index used by nbr_perm) with theirs[k] = mine[perm[p][k]] for the face
// vertices of two cells, matched by coordinates up to one common translation: the identity
// translation on a shared face, a period on a periodic one. -1 when none matches.
int periodic_face_perm(const HostMesh& m, int64_t e, int f, int64_t en, int fn) {
  static const int perms2[2][3] = {{0, 1, 0}, {1, 0, 0}};
  static const int perms3[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  const int d = m.dim, nv = d + 1, nperm = d == 2 ? 2 : 6;
  const int* fv = face_vertices(d);
  double mine[3][3], theirs[3][3];
  double scale = 0;
  for (int k = 0; k < d; ++k)
    for (int x = 0; x < d; ++x) {
      mine[k][x] = m.coords[(size_t)m.cell2nodes[(size_t)e * nv + fv[f * d + k]] * d + x];
      theirs[k][x] = m.coords[(size_t)m.cell2nodes[(size_t)en * nv + fv[fn * d + k]] * d + x];
    }
  for (int x = 0; x < d; ++x) scale = std::max(scale, std::fabs(mine[1][x] - mine[0][x]));
  const double tol = 1e-9 * scale;
  for (int p = 0; p < nperm; ++p) {
    const int* src = d == 2 ? perms2[p] : perms3[p];
    double t[3];
    for (int x = 0; x < d; ++x) t[x] = theirs[0][x] - mine[src[0]][x];
    bool ok = true;
    for (int k = 1; k < d && ok; ++k)
      for (int x = 0; x < d; ++x) ok &= std::fabs(theirs[k][x] - mine[src[k]][x] - t[x]) <= tol;
    if (ok) return p;
  }
  return -1;
}
}  // namespace

void adopt_op2_faces(HostMesh& m, int64_t n_faces, const int* face2cells, const int* faceNum,
                     int64_t n_bfaces, const int* bface2cells, const int* bfaceNum) {
  const int Nf = m.dim + 1;
  auto bad = [](const std::string& what) { throw std::invalid_argument("face maps: " + what); };
  if (n_faces < 0 || n_bfaces < 0) bad("negative set size");
  if ((n_faces > 0 && (!face2cells || !faceNum)) || (n_bfaces > 0 && (!bface2cells || !bfaceNum)))
    bad("null map");
  std::vector<char> seen((size_t)m.n_cells * Nf, 0);
  auto slot = [&](int c, int f) -> size_t {
    if (c < 0 || c >= m.n_cells || f < 0 || f >= Nf) bad("cell or local face out of range");
    return (size_t)c * Nf + f;
  };
  int64_t periodic = 0;
  for (int64_t i = 0; i < n_faces; ++i) {
    const int c0 = face2cells[2 * i], c1 = face2cells[2 * i + 1];
    const int f0 = faceNum[2 * i], f1 = faceNum[2 * i + 1];
    const size_t s0 = slot(c0, f0), s1 = slot(c1, f1);
    if (seen[s0] || seen[s1]) bad("face " + std::to_string(i) + " given twice");
    seen[s0] = seen[s1] = 1;
    if (m.nbr[s0] == c1 && m.nbr_face[s0] == f1) continue;
    // a periodic face: both sides lie on the mesh boundary and match by a translation
    const int p01 = m.nbr[s0] < 0 && m.nbr[s1] < 0 && c0 != c1 ? periodic_face_perm(m, c0, f0, c1, f1) : -1;
    const int p10 = p01 >= 0 ? periodic_face_perm(m, c1, f1, c0, f0) : -1;
    if (p10 < 0)
      bad("face " + std::to_string(i) + " (cells " + std::to_string(c0) + ", " + std::to_string(c1) +
          ", local faces " + std::to_string(f0) + ", " + std::to_string(f1) +
          ") is neither a face of the mesh with those local face numbers nor a periodic pair");
    m.nbr[s0] = c1;
    m.nbr_face[s0] = f1;
    m.nbr_perm[s0] = p01;
    m.nbr[s1] = c0;
    m.nbr_face[s1] = f0;
    m.nbr_perm[s1] = p10;
    ++periodic;
  }
  const int64_t interior = m.n_faces + periodic, boundary = m.n_bfaces - 2 * periodic;
  if (n_faces != interior)
    bad("n_faces " + std::to_string(n_faces) + " but the mesh has " + std::to_string(interior) +
        " interior faces");
  if (n_bfaces != boundary)
    bad("n_bfaces " + std::to_string(n_bfaces) + " but the mesh has " + std::to_string(boundary) +
        " boundary faces");
  for (int64_t b = 0; b < n_bfaces; ++b) {
    const size_t sb = slot(bface2cells[b], bfaceNum[b]);
    if (m.nbr[sb] >= 0)
      bad("bface " + std::to_string(b) + " (cell " + std::to_string(bface2cells[b]) + ", local face " +
          std::to_string(bfaceNum[b]) + ") is not on the boundary");
    if (seen[sb]) bad("bface " + std::to_string(b) + " given twice");
    seen[sb] = 1;
  }
  m.n_faces = n_faces;
  m.n_bfaces = n_bfaces;
  m.n_periodic = periodic;
  m.face2cells.assign(face2cells, face2cells + 2 * n_faces);
  m.faceNum.assign(faceNum, faceNum + 2 * n_faces);
  m.bface2cells.assign(bface2cells, bface2cells + n_bfaces);
  m.bfaceNum.assign(bfaceNum, bfaceNum + n_bfaces);
  for (int64_t b = 0; b < n_bfaces; ++b)
    m.nbr[(size_t)bface2cells[b] * Nf + bfaceNum[b]] = -(int)b - 1;
}

namespace {
// Affine map of cell e: J = det(dx/dxi) and G[a][x] = d xi_a / d x.
void cell_metric(const HostMesh& m, int64_t e, double G[3][3], double& J) {
  const int d = m.dim, nv = d + 1;
  const int* c = &m.cell2nodes[(size_t)e * nv];
  double xr[3][3];  // xr[ref dir][phys coord]
  for (int a = 0; a < d; ++a)
    for (int x = 0; x < d; ++x)
      xr[a][x] = 0.5 * (m.coords[(size_t)c[a + 1] * d + x] - m.coords[(size_t)c[0] * d + x]);
  if (d == 2) {
    J = xr[0][0] * xr[1][1] - xr[1][0] * xr[0][1];
    G[0][0] = xr[1][1] / J;
    G[0][1] = -xr[1][0] / J;
    G[1][0] = -xr[0][1] / J;
    G[1][1] = xr[0][0] / J;
  } else {
    auto cross = [](const double* a, const double* b, double* o) {
      o[0] = a[1] * b[2] - a[2] * b[1];
      o[1] = a[2] * b[0] - a[0] * b[2];
      o[2] = a[0] * b[1] - a[1] * b[0];
    };
    double c12[3], c20[3], c01[3];
    cross(xr[1], xr[2], c12);
    cross(xr[2], xr[0], c20);
    cross(xr[0], xr[1], c01);
    J = xr[0][0] * c12[0] + xr[0][1] * c12[1] + xr[0][2] * c12[2];
    for (int x = 0; x < 3; ++x) {
      G[0][x] = c12[x] / J;
      G[1][x] = c20[x] / J;
      G[2][x] = c01[x] / J;
    }
  }
}
}  // namespace

void mesh_inverse_jacobians(const HostMesh& m, double* G) {
  const int d = m.dim;
  for (int64_t e = 0; e < m.n_cells; ++e) {
    double g[3][3], J;
    cell_metric(m, e, g, J);
    for (int a = 0; a < d; ++a)
      for (int x = 0; x < d; ++x) G[((size_t)e * d + a) * d + x] = g[a][x];
  }
}

void build_geometry(HostMesh& m) {
  const int d = m.dim, nv = d + 1;
  const int nk = d == 2 ? 3 : 6;
  m.J.resize(m.n_cells);
  m.Kt.resize((size_t)m.n_cells * nk);
  for (int64_t e = 0; e < m.n_cells; ++e) {
    const int* c = &m.cell2nodes[(size_t)e * nv];
    double xr[3][3];  // xr[ref dir][phys coord]
    for (int a = 0; a < d; ++a)
      for (int x = 0; x < d; ++x)
        xr[a][x] = 0.5 * (m.coords[(size_t)c[a + 1] * d + x] - m.coords[(size_t)c[0] * d + x]);
    double G[3][3];  // G[ref dir][phys coord] = d xi_a / d x
    double J;
    if (d == 2) {
      J = xr[0][0] * xr[1][1] - xr[1][0] * xr[0][1];
      G[0][0] = xr[1][1] / J;
      G[0][1] = -xr[1][0] / J;
      G[1][0] = -xr[0][1] / J;
      G[1][1] = xr[0][0] / J;
    } else {
      auto cross = [](const double* a, const double* b, double* o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
      };
      double c12[3], c20[3], c01[3];
      cross(xr[1], xr[2], c12);
      cross(xr[2], xr[0], c20);
      cross(xr[0], xr[1], c01);
      J = xr[0][0] * c12[0] + xr[0][1] * c12[1] + xr[0][2] * c12[2];
      for (int x = 0; x < 3; ++x) {
        G[0][x] = c12[x] / J;
        G[1][x] = c20[x] / J;
        G[2][x] = c01[x] / J;
      }
    }
    if (!(J > 0)) throw std::runtime_error("non-positive element Jacobian");
    m.J[e] = J;
    auto dotg = [&](int a, int b) {
      double s = 0;
      for (int x = 0; x < d; ++x) s += G[a][x] * G[b][x];
      return s * J;
    };
    double* K = &m.Kt[(size_t)e * nk];
    if (d == 2) {
      K[0] = dotg(0, 0); K[1] = dotg(0, 1); K[2] = dotg(1, 1);
    } else {
      K[0] = dotg(0, 0); K[1] = dotg(0, 1); K[2] = dotg(0, 2);
      K[3] = dotg(1, 1); K[4] = dotg(1, 2); K[5] = dotg(2, 2);
    }
  }
}

void mesh_node_coords(const HostMesh& m, const RefElement& re, double* xyz) {
  const int d = m.dim, Np = re.Np;
  const int64_t E = m.n_cells;
  for (int64_t e = 0; e < E; ++e) {
    const int* c = &m.cell2nodes[(size_t)e * (d + 1)];
    for (int i = 0; i < Np; ++i) {
      // affine map from the reference simplex: barycentric weights of the vertices
      double lam[4];
      const double r = re.r[i], s = re.s[i], t = d == 3 ? re.t[i] : 0.0;
      if (d == 2) {
        lam[0] = -(r + s) / 2; lam[1] = (1 + r) / 2; lam[2] = (1 + s) / 2;
      } else {
        lam[0] = -(1 + r + s + t) / 2; lam[1] = (1 + r) / 2; lam[2] = (1 + s) / 2; lam[3] = (1 + t) / 2;
      }
      for (int x = 0; x < d; ++x) {
        double v = 0;
        for (int k = 0; k <= d; ++k) v += lam[k] * m.coords[(size_t)c[k] * d + x];
        xyz[((size_t)x * E + e) * Np + i] = v;
      }
    }
  }
}

std::vector<int64_t> shape_class_order(const HostMesh& m, int mode) {
  const int d = m.dim, nv = d + 1;
  const int64_t E = m.n_cells;
  std::vector<int64_t> key((size_t)E), perm((size_t)E);
  double vol = 0;  // total measure
  for (int64_t e = 0; e < E; ++e) {
    // shape class: the edge vectors from the first vertex, each component rounded against
    // the largest one (-1, 0, 1); congruent cells of a structured subdivision share it
    const int* c = &m.cell2nodes[(size_t)e * nv];
    double comp[3][3] = {}, mx = 0;
    for (int a = 1; a < nv; ++a)
      for (int x = 0; x < d; ++x) {
        comp[a - 1][x] = m.coords[(size_t)c[a] * d + x] - m.coords[(size_t)c[0] * d + x];
        mx = std::max(mx, std::fabs(comp[a - 1][x]));
      }
    int64_t shape = 0;
    for (int a = 0; a < d; ++a)
      for (int x = 0; x < d; ++x) shape = shape * 3 + (int64_t)std::lround(comp[a][x] / mx) + 1;
    key[(size_t)e] = shape;
    perm[(size_t)e] = e;
    vol += d == 2 ? std::fabs(comp[0][0] * comp[1][1] - comp[0][1] * comp[1][0]) / 2
                  : std::fabs(comp[0][0] * (comp[1][1] * comp[2][2] - comp[1][2] * comp[2][1]) -
                              comp[0][1] * (comp[1][0] * comp[2][2] - comp[1][2] * comp[2][0]) +
                              comp[0][2] * (comp[1][0] * comp[2][1] - comp[1][1] * comp[2][0])) / 6;
  }
  std::vector<int64_t> classes(key);
  std::sort(classes.begin(), classes.end());
  const int64_t K = std::unique(classes.begin(), classes.end()) - classes.begin();
  const auto by_class = [&](int64_t a, int64_t b) { return key[(size_t)a] < key[(size_t)b]; };
  if (mode == 2 && K <= 32 && E > 0) {
    // tiled: cubes (squares) of side s holding about 64 cells of each class, s^d = 64 K |cell|;
    // tiles in lexicographic order, inside a tile by class and old order. On the Kuhn cube a
    // tile is 4^3 hexes, a wave one class of one tile, and most neighbours share the tile.
    const double side = std::pow(64.0 * (double)K * vol / (double)E, 1.0 / d);
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int x = 0; x < d; ++x) lo[x] = hi[x] = m.coords[(size_t)x];
    for (int64_t v = 0; v < m.n_nodes; ++v)
      for (int x = 0; x < d; ++x) {
        lo[x] = std::min(lo[x], m.coords[(size_t)v * d + x]);
        hi[x] = std::max(hi[x], m.coords[(size_t)v * d + x]);
      }
    int64_t nt[3] = {1, 1, 1};
    for (int x = 0; x < d; ++x) nt[x] = (int64_t)std::floor((hi[x] - lo[x]) / side) + 1;
    std::vector<int64_t> tile((size_t)E);
    for (int64_t e = 0; e < E; ++e) {
      const int* c = &m.cell2nodes[(size_t)e * nv];
      int64_t t = 0;
      for (int x = d - 1; x >= 0; --x) {
        double cx = 0;
        for (int a = 0; a < nv; ++a) cx += m.coords[(size_t)c[a] * d + x];
        const int64_t i = std::min<int64_t>(nt[x] - 1, std::max<int64_t>(0, (int64_t)std::floor((cx / nv - lo[x]) / side)));
        t = t * nt[x] + i;
      }
      tile[(size_t)e] = t;
    }
    std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) {
      return tile[(size_t)a] != tile[(size_t)b] ? tile[(size_t)a] < tile[(size_t)b]
                                                : key[(size_t)a] < key[(size_t)b];
    });
    return perm;
  }
  // windowed: sort within runs of 64 K cells (K classes), so a wave of 64 cells still holds one
  // class on a structured subdivision while a cell's neighbours stay a few windows away;
  // with many classes (an unstructured mesh) the whole mesh is one window
  int64_t window = E;
  if (mode == 1 && K <= 32) window = 64 * K;
  for (int64_t b = 0; b < E; b += window)
    std::stable_sort(perm.begin() + b, perm.begin() + std::min(E, b + window), by_class);
  return perm;
}

HostMesh renumber_cells(const HostMesh& m, const std::vector<int64_t>& perm) {
  if ((int64_t)perm.size() != m.n_cells) throw std::invalid_argument("permutation size");
  if (m.n_periodic > 0)
    throw std::invalid_argument("meshes with periodic faces keep the caller's cell numbering");
  const int nv = m.dim + 1;
  std::vector<char> seen((size_t)m.n_cells, 0);
  HostMesh r;
  r.dim = m.dim;
  r.n_nodes = m.n_nodes;
  r.n_cells = m.n_cells;
  r.coords = m.coords;
  r.cell2nodes.resize(m.cell2nodes.size());
  for (int64_t i = 0; i < m.n_cells; ++i) {
    const int64_t o = perm[(size_t)i];
    if (o < 0 || o >= m.n_cells || seen[(size_t)o]) throw std::invalid_argument("not a permutation");
    seen[(size_t)o] = 1;
    for (int k = 0; k < nv; ++k) r.cell2nodes[(size_t)i * nv + k] = m.cell2nodes[(size_t)o * nv + k];
  }
  build_topology(r);
  build_geometry(r);
  return r;
}

}  // namespace insp
