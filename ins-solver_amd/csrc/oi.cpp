#include "oi.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>

namespace insp {

namespace {

// Fully symmetric triangle rules of degree 2N (N = 1..4): orbits S3 (centroid, weight),
// S21 (a, weight) = permutations of (a, a, 1-2a), S111 (a, b, weight) = permutations of
// (a, b, 1-a-b). Weights sum to 1. Dunavant (1985) rules 2, 4, 6, 8, refined by Newton on the
// moment equations to double precision.
struct Orbit {
  int kind;  // 0 S3, 1 S21, 2 S111
  double a, b, w;
};
const std::vector<Orbit>& tri_orbits(int N) {
  static const std::vector<Orbit> r1 = {{1, 0.16666666666666669, 0, 0.33333333333333331}};
  static const std::vector<Orbit> r2 = {{1, 0.44594849091596483, 0, 0.22338158967801119},
                                        {1, 0.091576213509770896, 0, 0.10995174365532209}};
  static const std::vector<Orbit> r3 = {{1, 0.24928674517091104, 0, 0.11678627572637827},
                                        {1, 0.063089014491501907, 0, 0.050844906370206382},
                                        {2, 0.053145049844817507, 0.31035245103378351, 0.082851075618374334}};
  static const std::vector<Orbit> r4 = {{0, 0, 0, 0.14431560767778651},
                                        {1, 0.45929258829272362, 0, 0.095091634267283495},
                                        {1, 0.17056930775176077, 0, 0.10321737053472199},
                                        {1, 0.05054722831703154, 0, 0.032458497623198745},
                                        {2, 0.0083947774099542016, 0.26311282963464122, 0.02723031417443347}};
  switch (N) {
    case 1: return r1;
    case 2: return r2;
    case 3: return r3;
    case 4: return r4;
  }
  throw std::invalid_argument("over-integration rules exist for orders 1..4");
}

// Barycentric points (3 per point) and weights summing to 1.
void tri_rule(int N, std::vector<double>& bary, std::vector<double>& w) {
  bary.clear();
  w.clear();
  for (const Orbit& o : tri_orbits(N)) {
    std::vector<std::array<double, 3>> pts;
    if (o.kind == 0) {
      pts.push_back({1.0 / 3, 1.0 / 3, 1.0 / 3});
    } else if (o.kind == 1) {
      const double c = 1 - 2 * o.a;
      pts = {{o.a, o.a, c}, {o.a, c, o.a}, {c, o.a, o.a}};
    } else {
      const double c = 1 - o.a - o.b;
      pts = {{o.a, o.b, c}, {o.a, c, o.b}, {o.b, o.a, c}, {o.b, c, o.a}, {c, o.a, o.b}, {c, o.b, o.a}};
    }
    for (const auto& p : pts) {
      bary.insert(bary.end(), p.begin(), p.end());
      w.push_back(o.w);
    }
  }
}

// Gauss-Legendre on (-1, 1) by Newton on the Legendre recurrence.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const double pi = std::acos(-1.0);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(pi * (i + 0.75) / (n + 0.5)), dp = 0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1, p1 = z;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[n - 1 - i] = z;
    w[n - 1 - i] = 2 / ((1 - z * z) * dp * dp);
  }
}

}  // namespace

OiRule make_oi_rule(int dim, int N) {
  if (N < 1 || N > kMaxOrder) throw std::invalid_argument("over-integration rules exist for orders 1..4");
  OiRule q;
  q.dim = dim;
  q.N = N;
  // reference simplex vertices (refelem.h conventions)
  const double V2[3][2] = {{-1, -1}, {1, -1}, {-1, 1}};
  const double V3[4][3] = {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
  std::vector<double> tb, tw;
  tri_rule(N, tb, tw);
  if (dim == 2) {
    // volume: the symmetric triangle rule, scaled to the reference area 2
    q.nv = (int)tw.size();
    for (int i = 0; i < q.nv; ++i) {
      for (int x = 0; x < 2; ++x) {
        double a = 0;
        for (int k = 0; k < 3; ++k) a += tb[i * 3 + k] * V2[k][x];
        q.vx.push_back(a);
      }
      q.vw.push_back(2.0 * tw[i]);
    }
    // faces: Gauss-Legendre with 2N+1 points, barycentrics (1-t, t) of the edge
    std::vector<double> gx, gw;
    gauss_legendre(2 * N + 1, gx, gw);
    q.ns = (int)gx.size();
    for (int i = 0; i < q.ns; ++i) {
      const double t = 0.5 * (gx[i] + 1);
      q.sb.push_back(1 - t);
      q.sb.push_back(t);
      q.sw.push_back(gw[i]);
    }
  } else {
    // volume: conical product of Gauss-Legendre rules (Duffy map), N+2 points per direction
    std::vector<double> gx, gw;
    const int n = N + 2;
    gauss_legendre(n, gx, gw);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k) {
          const double a = 0.5 * (gx[i] + 1), b = 0.5 * (gx[j] + 1), c = 0.5 * (gx[k] + 1);
          // unit-simplex point y and weight (Jacobian (1-a)^2 (1-b))
          const double y0 = a, y1 = b * (1 - a), y2 = c * (1 - a) * (1 - b);
          const double wt = 0.125 * gw[i] * gw[j] * gw[k] * (1 - a) * (1 - a) * (1 - b);
          const double l[4] = {1 - y0 - y1 - y2, y0, y1, y2};
          for (int x = 0; x < 3; ++x) {
            double p = 0;
            for (int v = 0; v < 4; ++v) p += l[v] * V3[v][x];
            q.vx.push_back(p);
          }
          q.vw.push_back(8.0 * wt);  // unit simplex (1/6) -> reference (4/3)
        }
    q.nv = (int)q.vw.size();
    // faces: the symmetric triangle rule, scaled to the reference face measure 2
    q.ns = (int)tw.size();
    q.sb = tb;
    for (double v : tw) q.sw.push_back(2.0 * v);
  }
  // neighbour point under each face vertex permutation (refelem.h: the neighbour's face
  // vertex k is my face vertex perms[p][k])
  const RefElement re = make_ref_element(dim, 1);
  q.qmap.assign((size_t)re.nperm * q.ns, -1);
  for (int p = 0; p < re.nperm; ++p)
    for (int i = 0; i < q.ns; ++i) {
      double bn[3];
      for (int k = 0; k < dim; ++k) bn[k] = q.sb[(size_t)i * dim + re.perms[p * dim + k]];
      for (int j = 0; j < q.ns; ++j) {
        double err = 0;
        for (int k = 0; k < dim; ++k) err += std::fabs(q.sb[(size_t)j * dim + k] - bn[k]);
        if (err < 1e-12) q.qmap[(size_t)p * q.ns + i] = j;
      }
      if (q.qmap[(size_t)p * q.ns + i] < 0) throw std::runtime_error("face cubature rule is not symmetric");
    }
  return q;
}

OiLayout oi_layout(int dim, int N, const OiRule& q) {
  OiLayout L;
  L.dim = dim;
  L.Np = dim == 2 ? (N + 1) * (N + 2) / 2 : (N + 1) * (N + 2) * (N + 3) / 6;
  L.Nf = dim + 1;
  L.nv = q.nv;
  L.ns = q.ns;
  L.off_GV = 0;
  L.off_wv = L.off_GV + dim * q.nv * L.Np;
  L.off_IS = L.off_wv + q.nv;
  L.off_GS = L.off_IS + L.Nf * q.ns * L.Np;
  L.off_ws = L.off_GS + L.Nf * dim * q.ns * L.Np;
  L.size = L.off_ws + q.ns;
  return L;
}

void oi_interp(const RefElement& re, const OiRule& q, std::vector<double>& Iv, std::vector<double>& Is,
               std::vector<double>& xv, std::vector<double>& xs) {
  const int dim = re.dim, Np = re.Np, Nf = re.Nfaces;
  const int* fv = face_vertices(dim);
  const double V2[3][2] = {{-1, -1}, {1, -1}, {-1, 1}};
  const double V3[4][3] = {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
  xv = q.vx;
  xs.assign((size_t)Nf * q.ns * dim, 0.0);
  for (int f = 0; f < Nf; ++f)
    for (int i = 0; i < q.ns; ++i)
      for (int x = 0; x < dim; ++x) {
        double a = 0;
        for (int k = 0; k < dim; ++k)
          a += q.sb[(size_t)i * dim + k] * (dim == 2 ? V2[fv[f * dim + k]][x] : V3[fv[f * dim + k]][x]);
        xs[((size_t)f * q.ns + i) * dim + x] = a;
      }
  basis_at(re, xv, Iv, nullptr);
  basis_at(re, xs, Is, nullptr);
}

std::vector<double> make_oi_tables(const RefElement& re, const OiRule& q) {
  const int dim = re.dim, Np = re.Np, Nf = re.Nfaces;
  const OiLayout L = oi_layout(dim, re.N, q);
  std::vector<double> C((size_t)L.size, 0.0);
  std::vector<double> Iv, Is, xv, xs;
  oi_interp(re, q, Iv, Is, xv, xs);
  std::vector<double> GV[3], GSf[3], tmp;
  basis_at(re, xv, tmp, GV);
  basis_at(re, xs, tmp, GSf);
  for (int s = 0; s < dim; ++s)
    for (int i = 0; i < q.nv; ++i)
      for (int m = 0; m < Np; ++m) C[L.off_GV + ((size_t)s * q.nv + i) * Np + m] = GV[s][(size_t)i * Np + m];
  for (int i = 0; i < q.nv; ++i) C[L.off_wv + i] = q.vw[i];
  for (int f = 0; f < Nf; ++f)
    for (int i = 0; i < q.ns; ++i)
      for (int m = 0; m < Np; ++m) {
        const size_t row = (size_t)f * q.ns + i;
        C[L.off_IS + row * Np + m] = Is[row * Np + m];
        for (int s = 0; s < dim; ++s)
          C[L.off_GS + (((size_t)f * dim + s) * q.ns + i) * Np + m] = GSf[s][row * Np + m];
      }
  for (int i = 0; i < q.ns; ++i) C[L.off_ws + i] = q.sw[i];
  return C;
}

}  // namespace insp
