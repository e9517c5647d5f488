// Host check of the smoothed-aggregation hierarchy used below order 1.
//
// Builds the P1 stiffness matrix K of a jittered Kuhn-tet cube (or triangle
// square) from textbook barycentric gradients, plus a lumped-mass shift
// (A = K + c M_L), sets up the hierarchy with the product code and checks:
//   - R = P^T exactly and A_{k+1} = R A_k P (Galerkin) to rounding;
//   - coarsening ratio and operator complexity are in the usual SA range;
//   - CG preconditioned by one V-cycle (Chebyshev(2) on D^-1 A, interval
//     [lmax/20, 1.1 lmax], dense coarsest solve) reaches 1e-10 in few iterations,
//     and the count barely grows under refinement (h-robustness).
// Prints one line per mesh: "dim n rows levels iters opcomplex galerkin_err".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "amg.h"
#include "mesh.h"

using namespace insp;

static HostCsr stiffness(const HostMesh& m, double shift) {
  HostCsr A = vertex_graph(m);
  const int d = m.dim, d1 = d + 1;
  std::vector<double> dense;  // per-cell element matrix
  for (int64_t e = 0; e < m.n_cells; ++e) {
    const int* c = &m.cell2nodes[(size_t)e * d1];
    // Jacobian columns x_i - x_0
    double Jm[3][3] = {{0}};
    for (int i = 1; i <= d; ++i)
      for (int x = 0; x < d; ++x) Jm[x][i - 1] = m.coords[(size_t)c[i] * d + x] - m.coords[(size_t)c[0] * d + x];
    double det, inv[3][3];
    if (d == 2) {
      det = Jm[0][0] * Jm[1][1] - Jm[0][1] * Jm[1][0];
      inv[0][0] = Jm[1][1] / det; inv[0][1] = -Jm[0][1] / det;
      inv[1][0] = -Jm[1][0] / det; inv[1][1] = Jm[0][0] / det;
    } else {
      det = Jm[0][0] * (Jm[1][1] * Jm[2][2] - Jm[1][2] * Jm[2][1]) -
            Jm[0][1] * (Jm[1][0] * Jm[2][2] - Jm[1][2] * Jm[2][0]) +
            Jm[0][2] * (Jm[1][0] * Jm[2][1] - Jm[1][1] * Jm[2][0]);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          const int i1 = (j + 1) % 3, i2 = (j + 2) % 3, j1 = (i + 1) % 3, j2 = (i + 2) % 3;
          inv[i][j] = (Jm[i1][j1] * Jm[i2][j2] - Jm[i1][j2] * Jm[i2][j1]) / det;
        }
    }
    const double vol = std::fabs(det) / (d == 2 ? 2.0 : 6.0);
    // gradients of barycentric coordinates: rows of inv for lambda_1..d, lambda_0 = -sum
    double g[4][3] = {{0}};
    for (int i = 1; i <= d; ++i)
      for (int x = 0; x < d; ++x) g[i][x] = inv[i - 1][x];
    for (int x = 0; x < d; ++x) {
      double s = 0;
      for (int i = 1; i <= d; ++i) s += g[i][x];
      g[0][x] = -s;
    }
    for (int i = 0; i < d1; ++i)
      for (int j = 0; j < d1; ++j) {
        double v = 0;
        for (int x = 0; x < d; ++x) v += g[i][x] * g[j][x];
        v *= vol;
        if (i == j) v += shift * vol / d1;
        const int r = c[i], cc = c[j];
        for (int k = A.ptr[r]; k < A.ptr[r + 1]; ++k)
          if (A.col[k] == cc) A.val[k] += v;
      }
  }
  return A;
}

struct Cycle {
  const AmgHierarchy& H;
  std::vector<std::vector<double>> x, b, r, dd;
  explicit Cycle(const AmgHierarchy& h) : H(h) {
    for (auto& L : H.levels) {
      x.emplace_back(L.A.n);
      b.emplace_back(L.A.n);
      r.emplace_back(L.A.n);
      dd.emplace_back(L.A.n);
    }
  }
  void smooth(int k, bool zero) {
    const AmgHostLevel& L = H.levels[k];
    const int n = L.A.n;
    const double hi = 1.1 * L.lmax, lo = hi / 20.0;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    double rho_old = 1.0 / sigma;
    for (int j = 1; j <= 2; ++j) {
      if (zero && j == 1) std::fill(r[k].begin(), r[k].end(), 0.0);
      else csr_matvec(L.A, x[k].data(), r[k].data());
      double c1 = 0, c2 = 1.0 / theta;
      if (j > 1) {
        const double rho = 1.0 / (2.0 * sigma - rho_old);
        c1 = rho * rho_old;
        c2 = 2.0 * rho / delta;
        rho_old = rho;
      }
      for (int i = 0; i < n; ++i) {
        const double res = (b[k][i] - r[k][i]) / L.diag[i];
        dd[k][i] = c1 * dd[k][i] + c2 * res;
        x[k][i] = (zero && j == 1) ? dd[k][i] : x[k][i] + dd[k][i];
      }
    }
  }
  void run(int k) {
    const AmgHostLevel& L = H.levels[k];
    if (k + 1 == (int)H.levels.size()) {
      const int n = H.n_coarse;
      for (int i = 0; i < n; ++i) {
        double s = 0;
        for (int j = 0; j < n; ++j) s += H.coarse_inv[(size_t)i * n + j] * b[k][j];
        x[k][i] = s;
      }
      return;
    }
    smooth(k, true);
    csr_matvec(L.A, x[k].data(), r[k].data());
    for (int i = 0; i < L.A.n; ++i) r[k][i] = b[k][i] - r[k][i];
    csr_matvec(L.R, r[k].data(), b[k + 1].data());
    run(k + 1);
    std::vector<double> px(L.A.n);
    csr_matvec(L.P, x[k + 1].data(), px.data());
    for (int i = 0; i < L.A.n; ++i) x[k][i] += px[i];
    smooth(k, false);
  }
};

static double dotv(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

int main(int argc, char** argv) {
  const int dim = argc > 1 ? atoi(argv[1]) : 3;
  const int n = argc > 2 ? atoi(argv[2]) : 12;
  HostMesh m = dim == 2 ? make_square_mesh(n, n, 0.1, 7) : make_cube_mesh(n, 0.1, 7);
  build_topology(m);
  HostCsr A = stiffness(m, 1.0);
  const HostCsr A0 = A;
  AmgOptions opt;
  if (getenv("AMG_THETA")) opt.theta = atof(getenv("AMG_THETA"));
  AmgHierarchy H = amg_setup(A, opt, false);
  // Galerkin and transpose checks
  double gerr = 0, nnz = 0;
  for (size_t k = 0; k + 1 < H.levels.size(); ++k) {
    const AmgHostLevel& L = H.levels[k];
    HostCsr Pt = csr_transpose(L.P);
    if (Pt.ptr != L.R.ptr || Pt.col != L.R.col) gerr = 1e300;
    for (size_t q = 0; q < Pt.val.size() && gerr < 1e300; ++q) gerr = std::fmax(gerr, std::fabs(Pt.val[q] - L.R.val[q]));
    HostCsr G = csr_multiply(L.R, csr_multiply(L.A, L.P));
    const HostCsr& Ac = H.levels[k + 1].A;
    double amax = 0;
    for (double v : Ac.val) amax = std::fmax(amax, std::fabs(v));
    if (G.ptr != Ac.ptr || G.col != Ac.col) gerr = 1e300;
    for (size_t q = 0; q < G.val.size() && gerr < 1e300; ++q)
      gerr = std::fmax(gerr, std::fabs(G.val[q] - Ac.val[q]) / amax);
  }
  for (auto& L : H.levels) nnz += (double)L.A.nnz();
  const double opc = nnz / (double)A0.nnz();
  // PCG with one V-cycle
  const int N0 = A0.n;
  std::vector<double> b(N0), x(N0, 0.0), r(N0), z(N0), p(N0), Ap(N0);
  for (int i = 0; i < N0; ++i) b[i] = std::sin(0.37 * i) + 0.5;
  r = b;
  Cycle cyc(H);
  auto prec = [&](const std::vector<double>& rr, std::vector<double>& zz) {
    cyc.b[0] = rr;
    cyc.run(0);
    zz = cyc.x[0];
  };
  prec(r, z);
  p = z;
  double rz = dotv(r, z);
  const double bn = std::sqrt(dotv(b, b));
  int it = 0;
  for (; it < 200 && std::sqrt(dotv(r, r)) > 1e-10 * bn; ++it) {
    csr_matvec(A0, p.data(), Ap.data());
    const double a = rz / dotv(p, Ap);
    for (int i = 0; i < N0; ++i) {
      x[i] += a * p[i];
      r[i] -= a * Ap[i];
    }
    prec(r, z);
    const double rzn = dotv(r, z);
    for (int i = 0; i < N0; ++i) p[i] = z[i] + (rzn / rz) * p[i];
    rz = rzn;
  }
  printf("%d %d %d %zu %d %.3f %.3e\n", dim, n, N0, H.levels.size(), it, opc, gerr);
  if (getenv("AMG_VERBOSE"))
    for (auto& L : H.levels) printf("  level n=%d nnz/row=%.1f lmax=%.3f\n", L.A.n, (double)L.A.nnz() / L.A.n, L.lmax);
  return 0;
}
