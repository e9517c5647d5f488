// Reference-element construction: Warburton nodes, orthonormal simplex bases
// (Dubiner/Koornwinder), derivative, mass and face-mass matrices.
// Algorithms follow Hesthaven & Warburton (2008): JacobiP/JacobiGL (ch. 3),
// Nodes2D/Simplex2DP/GradSimplex2DP (ch. 6), Nodes3D/Simplex3DP/
// GradSimplex3DP (ch. 10). The reference's own tables come from its
// DGConstants (toolkit, not vendored); these are regenerated from the
// published construction.
#include "refelem.h"
#include "dg_layout.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace insp {

namespace {

const int kFaceVerts2D[3 * 2] = {0, 1, 1, 2, 2, 0};
const int kFaceVerts3D[4 * 3] = {0, 1, 2, 0, 1, 3, 1, 2, 3, 0, 2, 3};

// Orthonormal Jacobi polynomial P_n^{(a,b)} at x (H&W JacobiP).
double jacobi_p(double x, double a, double b, int n) {
  const double g0 = std::pow(2.0, a + b + 1) / (a + b + 1) * std::tgamma(a + 1) *
                    std::tgamma(b + 1) / std::tgamma(a + b + 1);
  double p0 = 1.0 / std::sqrt(g0);
  if (n == 0) return p0;
  const double g1 = (a + 1) * (b + 1) / (a + b + 3) * g0;
  double p1 = ((a + b + 2) * x / 2 + (a - b) / 2) / std::sqrt(g1);
  if (n == 1) return p1;
  double aold = 2.0 / (2 + a + b) * std::sqrt((a + 1) * (b + 1) / (a + b + 3));
  for (int i = 1; i < n; ++i) {
    const double h1 = 2 * i + a + b;
    const double anew = 2.0 / (h1 + 2) *
        std::sqrt((i + 1) * (i + 1 + a + b) * (i + 1 + a) * (i + 1 + b) / (h1 + 1) / (h1 + 3));
    const double bnew = -(a * a - b * b) / h1 / (h1 + 2);
    const double p2 = 1.0 / anew * (-aold * p0 + (x - bnew) * p1);
    aold = anew;
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

double grad_jacobi_p(double x, double a, double b, int n) {
  if (n == 0) return 0.0;
  return std::sqrt(n * (n + a + b + 1)) * jacobi_p(x, a + 1, b + 1, n - 1);
}

// Legendre-Gauss-Lobatto points on [-1,1], ascending: the end points and the
// roots of P^{(1,1)}_{N-1}, found by Newton iteration with deflation.
std::vector<double> gauss_lobatto(int N) {
  std::vector<double> x(N + 1);
  x[0] = -1.0;
  x[N] = 1.0;
  const int m = N - 1;
  std::vector<double> roots(m);
  for (int k = 0; k < m; ++k) roots[k] = -std::cos((2.0 * k + 1) * M_PI / (2.0 * m));
  for (int k = 0; k < m; ++k) {
    double z = roots[k];
    for (int it = 0; it < 100; ++it) {
      double f = jacobi_p(z, 1, 1, m), df = grad_jacobi_p(z, 1, 1, m);
      double defl = 0.0;
      for (int j = 0; j < k; ++j) defl += 1.0 / (z - roots[j]);
      const double dz = f / (df - f * defl);
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    roots[k] = z;
  }
  std::sort(roots.begin(), roots.end());
  for (int k = 0; k < m; ++k) x[k + 1] = roots[k];
  return x;
}

// 1D warp function evaluated at rout (H&W Warpfactor): interpolant of the
// GLL-minus-equidistant displacement, divided by the edge blend 1 - r^2.
double warp_factor(int N, double rout) {
  const std::vector<double> lgl = gauss_lobatto(N);
  // Lagrange interpolation through equidistant points of (lgl - req).
  double warp = 0.0;
  for (int i = 0; i <= N; ++i) {
    const double xi = -1.0 + 2.0 * i / N;
    double li = 1.0;
    for (int j = 0; j <= N; ++j) {
      if (j == i) continue;
      const double xj = -1.0 + 2.0 * j / N;
      li *= (rout - xj) / (xi - xj);
    }
    warp += li * (lgl[i] - xi);
  }
  if (std::fabs(rout) < 1.0 - 1e-10) warp /= (1.0 - rout * rout);
  else warp = 0.0;  // the blend vanishes on the end points
  return warp;
}

void nodes_2d(int N, std::vector<double>& r, std::vector<double>& s) {
  static const double alpopt[15] = {0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832,
                                    1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258};
  const double alpha = N < 16 ? alpopt[N - 1] : 5.0 / 3.0;
  const int Np = (N + 1) * (N + 2) / 2;
  r.assign(Np, 0.0);
  s.assign(Np, 0.0);
  int sk = 0;
  for (int n = 0; n <= N; ++n) {
    for (int m = 0; m <= N - n; ++m) {
      const double L1 = double(n) / N, L3 = double(m) / N, L2 = 1.0 - L1 - L3;
      double x = -L2 + L3, y = (-L2 - L3 + 2 * L1) / std::sqrt(3.0);
      const double b1 = 4 * L2 * L3, b2 = 4 * L1 * L3, b3 = 4 * L1 * L2;
      const double w1 = b1 * warp_factor(N, L3 - L2) * (1 + (alpha * L1) * (alpha * L1));
      const double w2 = b2 * warp_factor(N, L1 - L3) * (1 + (alpha * L2) * (alpha * L2));
      const double w3 = b3 * warp_factor(N, L2 - L1) * (1 + (alpha * L3) * (alpha * L3));
      x += w1 + std::cos(2 * M_PI / 3) * w2 + std::cos(4 * M_PI / 3) * w3;
      y += std::sin(2 * M_PI / 3) * w2 + std::sin(4 * M_PI / 3) * w3;
      // equilateral (x,y) -> reference (r,s)
      const double l1 = (std::sqrt(3.0) * y + 1.0) / 3.0;
      const double l2 = (-3.0 * x - std::sqrt(3.0) * y + 2.0) / 6.0;
      const double l3 = (3.0 * x - std::sqrt(3.0) * y + 2.0) / 6.0;
      r[sk] = -l2 + l3 - l1;
      s[sk] = -l2 - l3 + l1;
      ++sk;
    }
  }
}

// Warp on a face of the equilateral tetrahedron (H&W evalshift / evalwarp).
void eval_shift(int N, double pval, double L1, double L2, double L3, double& dx, double& dy) {
  const std::vector<double> gl = gauss_lobatto(N);
  auto evalwarp = [&](double xout) {
    // xnodes = -gl (descending), xeq descending equidistant.
    double warp = 0.0;
    for (int i = 0; i <= N; ++i) {
      const double xeq_i = -1.0 + 2.0 * (N - i) / N;
      double d = (-gl[i]) - xeq_i;
      for (int j = 1; j < N; ++j) {
        if (i == j) continue;
        const double xeq_j = -1.0 + 2.0 * (N - j) / N;
        d *= (xout - xeq_j) / (xeq_i - xeq_j);
      }
      if (i != 0) d = -d / (xeq_i - 1.0);
      if (i != N) d = d / (xeq_i + 1.0);
      warp += d;
    }
    return warp;
  };
  const double warp1 = evalwarp(L3 - L2), warp2 = evalwarp(L1 - L3), warp3 = evalwarp(L2 - L1);
  const double wf1 = 4 * L2 * L3 * warp1 * (1 + (pval * L1) * (pval * L1));
  const double wf2 = 4 * L1 * L3 * warp2 * (1 + (pval * L2) * (pval * L2));
  const double wf3 = 4 * L1 * L2 * warp3 * (1 + (pval * L3) * (pval * L3));
  dx = wf1 + std::cos(2 * M_PI / 3) * wf2 + std::cos(4 * M_PI / 3) * wf3;
  dy = std::sin(2 * M_PI / 3) * wf2 + std::sin(4 * M_PI / 3) * wf3;
}

void nodes_3d(int N, std::vector<double>& r, std::vector<double>& s, std::vector<double>& t) {
  static const double alpopt[15] = {0, 0, 0, 0.1002, 1.1332, 1.5608, 1.3413, 1.2577,
                                    1.1603, 1.10153, 0.6080, 0.4523, 0.8856, 0.8717, 0.9655};
  const double alpha = N <= 15 ? alpopt[N - 1] : 1.0;
  const double tol = 1e-10;
  const int Np = (N + 1) * (N + 2) * (N + 3) / 6;
  r.assign(Np, 0.0);
  s.assign(Np, 0.0);
  t.assign(Np, 0.0);
  const double v1[3] = {-1, -1 / std::sqrt(3.0), -1 / std::sqrt(6.0)};
  const double v2[3] = {1, -1 / std::sqrt(3.0), -1 / std::sqrt(6.0)};
  const double v3[3] = {0, 2 / std::sqrt(3.0), -1 / std::sqrt(6.0)};
  const double v4[3] = {0, 0, 3 / std::sqrt(6.0)};
  double t1[4][3], t2[4][3];
  for (int k = 0; k < 3; ++k) {
    t1[0][k] = v2[k] - v1[k];
    t1[1][k] = v2[k] - v1[k];
    t1[2][k] = v3[k] - v2[k];
    t1[3][k] = v3[k] - v1[k];
    t2[0][k] = v3[k] - 0.5 * (v1[k] + v2[k]);
    t2[1][k] = v4[k] - 0.5 * (v1[k] + v2[k]);
    t2[2][k] = v4[k] - 0.5 * (v2[k] + v3[k]);
    t2[3][k] = v4[k] - 0.5 * (v1[k] + v3[k]);
  }
  for (int f = 0; f < 4; ++f) {
    double n1 = 0, n2 = 0;
    for (int k = 0; k < 3; ++k) {
      n1 += t1[f][k] * t1[f][k];
      n2 += t2[f][k] * t2[f][k];
    }
    for (int k = 0; k < 3; ++k) {
      t1[f][k] /= std::sqrt(n1);
      t2[f][k] /= std::sqrt(n2);
    }
  }
  int sk = 0;
  for (int n = 0; n <= N; ++n) {
    for (int m = 0; m <= N - n; ++m) {
      for (int q = 0; q <= N - n - m; ++q) {
        const double rr = -1 + 2.0 * q / N, ss = -1 + 2.0 * m / N, tt = -1 + 2.0 * n / N;
        const double L1 = (1 + tt) / 2, L2 = (1 + ss) / 2, L3 = -(1 + rr + ss + tt) / 2,
                     L4 = (1 + rr) / 2;
        double X[3];
        for (int k = 0; k < 3; ++k) X[k] = L3 * v1[k] + L4 * v2[k] + L2 * v3[k] + L1 * v4[k];
        double shift[3] = {0, 0, 0};
        for (int f = 0; f < 4; ++f) {
          double La, Lb, Lc, Ld;
          if (f == 0) { La = L1; Lb = L2; Lc = L3; Ld = L4; }
          else if (f == 1) { La = L2; Lb = L1; Lc = L3; Ld = L4; }
          else if (f == 2) { La = L3; Lb = L1; Lc = L4; Ld = L2; }
          else { La = L4; Lb = L1; Lc = L3; Ld = L2; }
          double w1, w2;
          eval_shift(N, alpha, Lb, Lc, Ld, w1, w2);
          double blend = Lb * Lc * Ld;
          const double denom = (Lb + 0.5 * La) * (Lc + 0.5 * La) * (Ld + 0.5 * La);
          if (denom > tol) blend = (1 + (alpha * La) * (alpha * La)) * blend / denom;
          const int nz = (Lb > tol) + (Lc > tol) + (Ld > tol);
          if (La < tol && nz < 3) {
            for (int k = 0; k < 3; ++k) shift[k] = w1 * t1[f][k] + w2 * t2[f][k];
          } else {
            for (int k = 0; k < 3; ++k) shift[k] += blend * w1 * t1[f][k] + blend * w2 * t2[f][k];
          }
        }
        for (int k = 0; k < 3; ++k) X[k] += shift[k];
        // equilateral -> reference: X - 0.5*(v2+v3+v4-v1) = A * (r,s,t)
        double rhs[3], A[3][3];
        for (int k = 0; k < 3; ++k) {
          rhs[k] = X[k] - 0.5 * (v2[k] + v3[k] + v4[k] - v1[k]);
          A[k][0] = 0.5 * (v2[k] - v1[k]);
          A[k][1] = 0.5 * (v3[k] - v1[k]);
          A[k][2] = 0.5 * (v4[k] - v1[k]);
        }
        std::vector<double> Am(9);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) Am[a * 3 + b] = A[a][b];
        const std::vector<double> Ai = mat_inverse(Am, 3);
        double rst[3];
        for (int a = 0; a < 3; ++a) rst[a] = Ai[a * 3] * rhs[0] + Ai[a * 3 + 1] * rhs[1] + Ai[a * 3 + 2] * rhs[2];
        r[sk] = rst[0];
        s[sk] = rst[1];
        t[sk] = rst[2];
        ++sk;
      }
    }
  }
}

// Orthonormal basis on the triangle and its (r,s) gradient.
void simplex2d(double r, double s, int i, int j, double& P, double& Pr, double& Ps) {
  const double a = (std::fabs(s - 1.0) > 1e-14) ? 2 * (1 + r) / (1 - s) - 1 : -1.0;
  const double b = s;
  const double fa = jacobi_p(a, 0, 0, i), dfa = grad_jacobi_p(a, 0, 0, i);
  const double gb = jacobi_p(b, 2 * i + 1, 0, j), dgb = grad_jacobi_p(b, 2 * i + 1, 0, j);
  P = std::sqrt(2.0) * fa * gb * std::pow(1 - b, i);
  double dr = dfa * gb;
  if (i > 0) dr *= std::pow(0.5 * (1 - b), i - 1);
  double ds = dfa * (gb * (0.5 * (1 + a)));
  if (i > 0) ds *= std::pow(0.5 * (1 - b), i - 1);
  double tmp = dgb * std::pow(0.5 * (1 - b), i);
  if (i > 0) tmp -= 0.5 * i * gb * std::pow(0.5 * (1 - b), i - 1);
  ds += fa * tmp;
  const double scale = std::pow(2.0, i + 0.5);
  Pr = dr * scale;
  Ps = ds * scale;
}

// Orthonormal basis on the tetrahedron and its (r,s,t) gradient.
void simplex3d(double r, double s, double t, int i, int j, int k, double& P, double& Pr,
               double& Ps, double& Pt) {
  const double a = (std::fabs(s + t) > 1e-14) ? 2 * (1 + r) / (-s - t) - 1 : -1.0;
  const double b = (std::fabs(t - 1.0) > 1e-14) ? 2 * (1 + s) / (1 - t) - 1 : -1.0;
  const double c = t;
  const double fa = jacobi_p(a, 0, 0, i), dfa = grad_jacobi_p(a, 0, 0, i);
  const double gb = jacobi_p(b, 2 * i + 1, 0, j), dgb = grad_jacobi_p(b, 2 * i + 1, 0, j);
  const double hc = jacobi_p(c, 2 * (i + j) + 2, 0, k),
               dhc = grad_jacobi_p(c, 2 * (i + j) + 2, 0, k);
  P = 2 * std::sqrt(2.0) * fa * gb * std::pow(1 - b, i) * hc * std::pow(1 - c, i + j);
  double vr = dfa * (gb * hc);
  if (i > 0) vr *= std::pow(0.5 * (1 - b), i - 1);
  if (i + j > 0) vr *= std::pow(0.5 * (1 - c), i + j - 1);
  double vs = 0.5 * (1 + a) * vr;
  double tmp = dgb * std::pow(0.5 * (1 - b), i);
  if (i > 0) tmp += -0.5 * i * (gb * std::pow(0.5 * (1 - b), i - 1));
  if (i + j > 0) tmp *= std::pow(0.5 * (1 - c), i + j - 1);
  tmp = fa * (tmp * hc);
  vs += tmp;
  double vt = 0.5 * (1 + a) * vr + 0.5 * (1 + b) * tmp;
  tmp = dhc * std::pow(0.5 * (1 - c), i + j);
  if (i + j > 0) tmp -= 0.5 * (i + j) * (hc * std::pow(0.5 * (1 - c), i + j - 1));
  tmp = fa * (gb * tmp);
  tmp *= std::pow(0.5 * (1 - b), i);
  vt += tmp;
  const double scale = std::pow(2.0, 2 * i + j + 1.5);
  Pr = vr * scale;
  Ps = vs * scale;
  Pt = vt * scale;
}

// Vandermonde matrices of the orthonormal basis at the given points.
void vandermonde(int dim, int N, const std::vector<double>& r, const std::vector<double>& s,
                 const std::vector<double>& t, std::vector<double>& V, std::vector<double>* Vd) {
  const int n = (int)r.size();
  const int Np = dim == 2 ? (N + 1) * (N + 2) / 2 : (N + 1) * (N + 2) * (N + 3) / 6;
  V.assign((size_t)n * Np, 0.0);
  if (Vd)
    for (int d = 0; d < dim; ++d) Vd[d].assign((size_t)n * Np, 0.0);
  for (int p = 0; p < n; ++p) {
    int col = 0;
    if (dim == 2) {
      for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N - i; ++j) {
          double P, Pr, Ps;
          simplex2d(r[p], s[p], i, j, P, Pr, Ps);
          V[(size_t)p * Np + col] = P;
          if (Vd) {
            Vd[0][(size_t)p * Np + col] = Pr;
            Vd[1][(size_t)p * Np + col] = Ps;
          }
          ++col;
        }
    } else {
      for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N - i; ++j)
          for (int k = 0; k <= N - i - j; ++k) {
            double P, Pr, Ps, Pt;
            simplex3d(r[p], s[p], t[p], i, j, k, P, Pr, Ps, Pt);
            V[(size_t)p * Np + col] = P;
            if (Vd) {
              Vd[0][(size_t)p * Np + col] = Pr;
              Vd[1][(size_t)p * Np + col] = Ps;
              Vd[2][(size_t)p * Np + col] = Pt;
            }
            ++col;
          }
    }
  }
}

// Barycentric coordinates of a reference point with respect to v0..v{dim}.
void barycentric(int dim, double r, double s, double t, double* lam) {
  if (dim == 2) {
    lam[0] = -(r + s) / 2;
    lam[1] = (1 + r) / 2;
    lam[2] = (1 + s) / 2;
  } else {
    lam[0] = -(1 + r + s + t) / 2;
    lam[1] = (1 + r) / 2;
    lam[2] = (1 + s) / 2;
    lam[3] = (1 + t) / 2;
  }
}

}  // namespace

const int* face_vertices(int dim) { return dim == 2 ? kFaceVerts2D : kFaceVerts3D; }

std::vector<double> mat_inverse(const std::vector<double>& A, int n) {
  std::vector<double> a(A), inv((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int i = c + 1; i < n; ++i)
      if (std::fabs(a[(size_t)i * n + c]) > std::fabs(a[(size_t)piv * n + c])) piv = i;
    if (std::fabs(a[(size_t)piv * n + c]) < 1e-300) throw std::runtime_error("singular matrix");
    if (piv != c)
      for (int j = 0; j < n; ++j) {
        std::swap(a[(size_t)c * n + j], a[(size_t)piv * n + j]);
        std::swap(inv[(size_t)c * n + j], inv[(size_t)piv * n + j]);
      }
    const double d = 1.0 / a[(size_t)c * n + c];
    for (int j = 0; j < n; ++j) {
      a[(size_t)c * n + j] *= d;
      inv[(size_t)c * n + j] *= d;
    }
    for (int i = 0; i < n; ++i) {
      if (i == c) continue;
      const double f = a[(size_t)i * n + c];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) {
        a[(size_t)i * n + j] -= f * a[(size_t)c * n + j];
        inv[(size_t)i * n + j] -= f * inv[(size_t)c * n + j];
      }
    }
  }
  return inv;
}

std::vector<double> mat_mul(const std::vector<double>& A, const std::vector<double>& B, int m,
                            int k, int n) {
  std::vector<double> C((size_t)m * n, 0.0);
  for (int i = 0; i < m; ++i)
    for (int l = 0; l < k; ++l) {
      const double a = A[(size_t)i * k + l];
      if (a == 0.0) continue;
      for (int j = 0; j < n; ++j) C[(size_t)i * n + j] += a * B[(size_t)l * n + j];
    }
  return C;
}

std::vector<double> mat_transpose(const std::vector<double>& A, int m, int n) {
  std::vector<double> T((size_t)m * n);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) T[(size_t)j * m + i] = A[(size_t)i * n + j];
  return T;
}

RefElement make_ref_element(int dim, int N) {
  if ((dim != 2 && dim != 3) || N < 1 || N > kMaxOrder)
    throw std::runtime_error("unsupported element (dim must be 2 or 3, order 1..4)");
  RefElement E;
  E.dim = dim;
  E.N = N;
  E.Nfaces = dim + 1;
  if (dim == 2) {
    E.Np = (N + 1) * (N + 2) / 2;
    E.Nfp = N + 1;
    nodes_2d(N, E.r, E.s);
    // N = 1 warp & blend places nodes exactly on the vertices; clean roundoff.
  } else {
    E.Np = (N + 1) * (N + 2) * (N + 3) / 6;
    E.Nfp = (N + 1) * (N + 2) / 2;
    nodes_3d(N, E.r, E.s, E.t);
  }
  const int Np = E.Np, Nfp = E.Nfp, Nf = E.Nfaces;
  std::vector<double> tt = dim == 3 ? E.t : std::vector<double>(Np, 0.0);
  std::vector<double> Vd[3];
  vandermonde(dim, N, E.r, E.s, tt, E.V, Vd);
  E.invV = mat_inverse(E.V, Np);
  for (int d = 0; d < dim; ++d) E.D[d] = mat_mul(Vd[d], E.invV, Np, Np, Np);
  const std::vector<double> VVt = mat_mul(E.V, mat_transpose(E.V, Np, Np), Np, Np, Np);
  E.invM = VVt;
  E.M = mat_inverse(VVt, Np);

  // Face masks: nodes whose barycentric coordinate opposite the face vanishes.
  const int* fv = face_vertices(dim);
  E.fmask.assign((size_t)Nf * Nfp, -1);
  E.face_bary.assign((size_t)Nf * Nfp * dim, 0.0);
  for (int f = 0; f < Nf; ++f) {
    // the vertex not on face f
    int opp = -1;
    for (int v = 0; v <= dim; ++v) {
      bool on = false;
      for (int k = 0; k < dim; ++k) on |= fv[f * dim + k] == v;
      if (!on) opp = v;
    }
    int cnt = 0;
    for (int p = 0; p < Np; ++p) {
      double lam[4];
      barycentric(dim, E.r[p], E.s[p], tt[p], lam);
      if (std::fabs(lam[opp]) < 1e-10) {
        if (cnt >= Nfp) throw std::runtime_error("face mask overflow");
        E.fmask[(size_t)f * Nfp + cnt] = p;
        for (int k = 0; k < dim; ++k)
          E.face_bary[((size_t)f * Nfp + cnt) * dim + k] = lam[fv[f * dim + k]];
        ++cnt;
      }
    }
    if (cnt != Nfp) throw std::runtime_error("face mask underflow");
  }

  // Face mass matrices in face coordinates (H&W Lift2D / Lift3D):
  // 2D faces use r, r, s; 3D faces use (r,s), (r,t), (s,t), (s,t).
  E.Mface.assign((size_t)Nf * Nfp * Nfp, 0.0);
  for (int f = 0; f < Nf; ++f) {
    std::vector<double> fr(Nfp), fs(Nfp);
    for (int j = 0; j < Nfp; ++j) {
      const int p = E.fmask[(size_t)f * Nfp + j];
      if (dim == 2) {
        fr[j] = (f == 2) ? E.s[p] : E.r[p];
      } else {
        if (f == 0) { fr[j] = E.r[p]; fs[j] = E.s[p]; }
        else if (f == 1) { fr[j] = E.r[p]; fs[j] = tt[p]; }
        else { fr[j] = E.s[p]; fs[j] = tt[p]; }
      }
    }
    std::vector<double> Vf((size_t)Nfp * Nfp);
    if (dim == 2) {
      for (int j = 0; j < Nfp; ++j)
        for (int i = 0; i <= N; ++i) Vf[(size_t)j * Nfp + i] = jacobi_p(fr[j], 0, 0, i);
    } else {
      std::vector<double> z(Nfp, 0.0);
      vandermonde(2, N, fr, fs, z, Vf, nullptr);
    }
    const std::vector<double> Mf =
        mat_inverse(mat_mul(Vf, mat_transpose(Vf, Nfp, Nfp), Nfp, Nfp, Nfp), Nfp);
    std::copy(Mf.begin(), Mf.end(), E.Mface.begin() + (size_t)f * Nfp * Nfp);
  }
  E.EMAT.assign((size_t)Np * Nf * Nfp, 0.0);
  for (int f = 0; f < Nf; ++f)
    for (int i = 0; i < Nfp; ++i)
      for (int j = 0; j < Nfp; ++j)
        E.EMAT[(size_t)E.fmask[(size_t)f * Nfp + i] * Nf * Nfp + f * Nfp + j] =
            E.Mface[((size_t)f * Nfp + i) * Nfp + j];

  // Face-node matching tables for every vertex permutation.
  if (dim == 2) {
    E.nperm = 2;
    E.perms = {0, 1, 1, 0};
  } else {
    E.nperm = 6;
    E.perms = {0, 1, 2, 0, 2, 1, 1, 0, 2, 1, 2, 0, 2, 0, 1, 2, 1, 0};
  }
  E.nbr_face_node.assign((size_t)Nf * Nf * E.nperm * Nfp, -1);
  for (int f = 0; f < Nf; ++f)
    for (int fn = 0; fn < Nf; ++fn)
      for (int p = 0; p < E.nperm; ++p)
        for (int j = 0; j < Nfp; ++j) {
          // my bary wrt my face vertices; neighbour vertex k == my vertex perms[p][k]
          const double* b = &E.face_bary[((size_t)f * Nfp + j) * dim];
          double bn[3];
          for (int k = 0; k < dim; ++k) bn[k] = b[E.perms[p * dim + k]];
          int found = -1;
          for (int q = 0; q < Nfp; ++q) {
            const double* c = &E.face_bary[((size_t)fn * Nfp + q) * dim];
            double err = 0;
            for (int k = 0; k < dim; ++k) err += std::fabs(c[k] - bn[k]);
            if (err < 1e-9) found = q;
          }
          if (found < 0) throw std::runtime_error("face node matching failed");
          E.nbr_face_node[(((size_t)f * Nf + fn) * E.nperm + p) * Nfp + j] = found;
        }
  return E;
}

void basis_at(const RefElement& re, const std::vector<double>& pts, std::vector<double>& I,
              std::vector<double>* G) {
  const int dim = re.dim, n = (int)(pts.size() / dim), Np = re.Np;
  std::vector<double> r(n), s(n), t(n, 0.0), V, Vd[3];
  for (int i = 0; i < n; ++i) {
    r[i] = pts[(size_t)i * dim];
    s[i] = pts[(size_t)i * dim + 1];
    if (dim == 3) t[i] = pts[(size_t)i * dim + 2];
  }
  vandermonde(dim, re.N, r, s, t, V, G ? Vd : nullptr);
  I = mat_mul(V, re.invV, n, Np, Np);
  if (G)
    for (int d = 0; d < dim; ++d) G[d] = mat_mul(Vd[d], re.invV, n, Np, Np);
}

std::vector<double> interp_matrix(const RefElement& to, const RefElement& from) {
  if (to.dim != from.dim) throw std::runtime_error("interp_matrix: dimension mismatch");
  std::vector<double> tt = to.dim == 3 ? to.t : std::vector<double>(to.Np, 0.0);
  std::vector<double> Vto;
  vandermonde(to.dim, from.N, to.r, to.s, tt, Vto, nullptr);  // to.Np x from.Np modal
  return mat_mul(Vto, from.invV, to.Np, from.Np, from.Np);
}

static P1NodeBits build_p1_node_bits(int dim) {
  const RefElement re = make_ref_element(dim, 1);
  const int Nf = dim + 1, Nfp = dim;
  P1NodeBits b{};
  for (int f = 0; f < Nf; ++f)
    for (int fn = 0; fn < Nf; ++fn)
      for (int p = 0; p < re.nperm; ++p)
        for (int j = 0; j < Nfp; ++j) {
          const int q = re.nbr_face_node[(((size_t)f * Nf + fn) * re.nperm + p) * Nfp + j];
          const unsigned long long node = (unsigned long long)re.fmask[fn * Nfp + q];
          b.w[f][j] |= node << (2 * (fn * re.nperm + p));
        }
  // in 3D the map does not depend on the cell's own face; the kernels read row 0 only
  if (dim == 3)
    for (int f = 1; f < Nf; ++f)
      for (int j = 0; j < Nfp; ++j)
        if (b.w[f][j] != b.w[0][j]) throw std::logic_error("order-1 node map depends on the face in 3D");
  return b;
}

const P1NodeBits& p1_node_bits(int dim) {
  static const P1NodeBits b2 = build_p1_node_bits(2), b3 = build_p1_node_bits(3);
  return dim == 2 ? b2 : b3;
}

}  // namespace insp
