// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only): the mesh
// generators, topology and geometry, cell renumbering, adoption of a caller's OP2 face sets,
// the OP2 HDF5 reader on the reference's mesh files, reference elements, the over-integration
// rules and their tables, and the AMG hierarchy set-up. Exits non-zero on a failed check; the
// sanitizers abort on any memory or UB error. Prints "ok" at the end.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "amg.h"
#include "hdf5_reader.h"
#include "mesh.h"
#include "oi.h"
#include "refelem.h"

using namespace insp;

#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      return 1;                                                    \
    }                                                              \
  } while (0)

static int mesh_checks(HostMesh m) {
  build_topology(m);
  build_geometry(m);
  CHECK(first_negative_cell(m) < 0);
  const int Nf = m.dim + 1;
  for (int64_t e = 0; e < m.n_cells; ++e)
    for (int f = 0; f < Nf; ++f) {
      const int n = m.nbr[(size_t)e * Nf + f];
      CHECK(n < m.n_cells && (n >= 0 || -n - 1 < m.n_bfaces));
    }
  for (int mode = 0; mode < 3; ++mode) {
    const std::vector<int64_t> perm = shape_class_order(m, mode);
    CHECK((int64_t)perm.size() == m.n_cells);
    HostMesh r = renumber_cells(m, perm);
    CHECK(r.n_cells == m.n_cells && r.n_bfaces == m.n_bfaces);
  }
  // the mesh's own faces given back as a caller's OP2 face sets, bfaces reversed
  HostMesh c = m;
  std::vector<int> b2c(m.bface2cells.rbegin(), m.bface2cells.rend());
  std::vector<int> bnum(m.bfaceNum.rbegin(), m.bfaceNum.rend());
  adopt_op2_faces(c, m.n_faces, m.face2cells.data(), m.faceNum.data(), m.n_bfaces, b2c.data(), bnum.data());
  CHECK(c.n_bfaces == m.n_bfaces && c.bface2cells == b2c);
  for (int N = 1; N <= kMaxOrder; ++N) {
    const RefElement re = make_ref_element(m.dim, N);
    std::vector<double> xyz((size_t)m.dim * m.n_cells * re.Np);
    mesh_node_coords(m, re, xyz.data());
    for (double v : xyz) CHECK(std::isfinite(v));
  }
  std::vector<double> G((size_t)m.n_cells * m.dim * m.dim);
  mesh_inverse_jacobians(m, G.data());
  return 0;
}

int main(int argc, char** argv) {
  if (mesh_checks(make_square_mesh(6, 5, 0.1, 3))) return 1;
  if (mesh_checks(make_cube_mesh(3, 0.1, 4))) return 1;
  // over-integration rules and tables, every dimension and order
  for (int d = 2; d <= 3; ++d)
    for (int N = 1; N <= kMaxOrder; ++N) {
      const OiRule q = make_oi_rule(d, N);
      const RefElement re = make_ref_element(d, N);
      const std::vector<double> C = make_oi_tables(re, q);
      CHECK((int)C.size() == oi_layout(d, N, q).size);
      double wsum = 0;
      for (double w : q.vw) wsum += w;
      CHECK(std::fabs(wsum - (d == 2 ? 2.0 : 4.0 / 3.0)) < 1e-13);
    }
  // OP2 HDF5 meshes given on the command line
  for (int i = 1; i < argc; ++i) {
    Hdf5File f(argv[i]);
    const std::vector<int32_t> c2n = f.read_int("cell2nodes");
    const std::vector<double> xy = f.read_double("node_coords");
    CHECK(!c2n.empty() && !xy.empty());
    for (const std::string& n : f.names()) CHECK(f.has(n));
  }
  // AMG hierarchy on the vertex graph of a small cube (structure only: a P1 Laplacian stand-in)
  {
    HostMesh m = make_cube_mesh(4, 0.1, 5);
    build_topology(m);
    HostCsr A = vertex_graph(m);
    A.val.assign((size_t)A.nnz(), 0.0);
    for (int r = 0; r < A.n; ++r)
      for (int k = A.ptr[r]; k < A.ptr[r + 1]; ++k)
        A.val[k] = A.col[k] == r ? (double)(A.ptr[r + 1] - A.ptr[r]) : -1.0 + 1e-3;
    const AmgHierarchy h = amg_setup(A, AmgOptions{}, false);
    CHECK(!h.levels.empty());
  }
  std::printf("ok\n");
  return 0;
}
