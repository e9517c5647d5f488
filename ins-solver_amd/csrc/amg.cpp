// Host side of the vertex coarse space: graph, colouring and the
// smoothed-aggregation hierarchy (Vanek, Mandel, Brezina 1996: strength of
// connection, three-phase greedy aggregation, tentative prolongator from the
// constant vector, one damped-Jacobi smoothing step, Galerkin coarse operators).
#include "amg.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <stdexcept>

#include "mesh.h"

namespace insp {

HostCsr csr_transpose(const HostCsr& A) {
  HostCsr T;
  T.n = A.m;
  T.m = A.n;
  T.ptr.assign((size_t)T.n + 1, 0);
  for (int k = 0; k < A.nnz(); ++k) ++T.ptr[(size_t)A.col[k] + 1];
  for (int i = 0; i < T.n; ++i) T.ptr[i + 1] += T.ptr[i];
  T.col.resize((size_t)A.nnz());
  T.val.resize((size_t)A.nnz());
  std::vector<int> pos(T.ptr.begin(), T.ptr.end() - 1);
  for (int i = 0; i < A.n; ++i)
    for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
      const int p = pos[A.col[k]]++;
      T.col[p] = i;
      T.val[p] = A.val[k];
    }
  return T;  // rows come out with sorted columns
}

HostCsr csr_multiply(const HostCsr& A, const HostCsr& B) {
  if (A.m != B.n) throw std::runtime_error("csr_multiply: shape mismatch");
  HostCsr C;
  C.n = A.n;
  C.m = B.m;
  C.ptr.assign((size_t)C.n + 1, 0);
  std::vector<int> mark((size_t)B.m, -1);
  std::vector<double> acc((size_t)B.m, 0.0);
  std::vector<int> cols;
  for (int i = 0; i < A.n; ++i) {
    cols.clear();
    for (int ka = A.ptr[i]; ka < A.ptr[i + 1]; ++ka) {
      const int j = A.col[ka];
      const double a = A.val[ka];
      for (int kb = B.ptr[j]; kb < B.ptr[j + 1]; ++kb) {
        const int c = B.col[kb];
        if (mark[c] != i) {
          mark[c] = i;
          acc[c] = 0.0;
          cols.push_back(c);
        }
        acc[c] += a * B.val[kb];
      }
    }
    std::sort(cols.begin(), cols.end());
    for (int c : cols) {
      C.col.push_back(c);
      C.val.push_back(acc[c]);
    }
    C.ptr[i + 1] = (int)C.col.size();
  }
  return C;
}

void csr_matvec(const HostCsr& A, const double* x, double* y) {
  for (int i = 0; i < A.n; ++i) {
    double s = 0.0;
    for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) s += A.val[k] * x[A.col[k]];
    y[i] = s;
  }
}

HostCsr vertex_graph(const HostMesh& m) {
  const int nv = (int)m.n_nodes, d1 = m.dim + 1;
  // vertex -> cells
  std::vector<int> cptr((size_t)nv + 1, 0);
  for (int64_t k = 0; k < m.n_cells * d1; ++k) ++cptr[(size_t)m.cell2nodes[k] + 1];
  for (int v = 0; v < nv; ++v) cptr[v + 1] += cptr[v];
  std::vector<int> cells((size_t)m.n_cells * d1), pos(cptr.begin(), cptr.end() - 1);
  for (int64_t e = 0; e < m.n_cells; ++e)
    for (int i = 0; i < d1; ++i) cells[pos[m.cell2nodes[e * d1 + i]]++] = (int)e;
  HostCsr g;
  g.n = g.m = nv;
  g.ptr.assign((size_t)nv + 1, 0);
  std::vector<int> mark((size_t)nv, -1), row;
  for (int v = 0; v < nv; ++v) {
    row.clear();
    for (int k = cptr[v]; k < cptr[v + 1]; ++k) {
      const int e = cells[k];
      for (int i = 0; i < d1; ++i) {
        const int w = m.cell2nodes[(size_t)e * d1 + i];
        if (mark[w] != v) {
          mark[w] = v;
          row.push_back(w);
        }
      }
    }
    if (mark[v] != v) row.push_back(v);  // a vertex in no cell still gets its diagonal
    std::sort(row.begin(), row.end());
    g.col.insert(g.col.end(), row.begin(), row.end());
    g.ptr[v + 1] = (int)g.col.size();
  }
  g.val.assign(g.col.size(), 0.0);
  return g;
}

std::vector<int> distance2_colouring(const HostCsr& g, int* n_colours) {
  const int n = g.n;
  std::vector<int> colour((size_t)n, -1), seen;  // seen[c] == v: colour c is taken near v
  int nc = 0;
  for (int v = 0; v < n; ++v) {
    for (int k = g.ptr[v]; k < g.ptr[v + 1]; ++k) {
      const int w = g.col[k];
      for (int q = g.ptr[w]; q < g.ptr[w + 1]; ++q) {
        const int c = colour[g.col[q]];
        if (c >= 0) seen[c] = v;
      }
    }
    int c = 0;
    while (c < nc && seen[c] == v) ++c;
    if (c == nc) {
      ++nc;
      seen.push_back(-1);
    }
    colour[v] = c;
  }
  *n_colours = nc;
  return colour;
}

namespace {

std::vector<double> diagonal(const HostCsr& A) {
  std::vector<double> d((size_t)A.n, 0.0);
  for (int i = 0; i < A.n; ++i)
    for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
      if (A.col[k] == i) d[i] = A.val[k];
  return d;
}

// lambda_max(D^-1 A) by power iteration from a seeded random start.
double lambda_max(const HostCsr& A, const std::vector<double>& d, int iters) {
  const int n = A.n;
  if (n == 0) return 1.0;
  std::mt19937_64 rng(0x5eedu + (uint64_t)n);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  std::vector<double> x((size_t)n), y((size_t)n);
  for (auto& v : x) v = U(rng);
  double lam = 1.0;
  for (int it = 0; it < iters; ++it) {
    csr_matvec(A, x.data(), y.data());
    double num = 0, den = 0, nrm = 0;
    for (int i = 0; i < n; ++i) {
      y[i] = d[i] != 0 ? y[i] / d[i] : 0.0;
      num += x[i] * y[i];
      den += x[i] * x[i];
      nrm += y[i] * y[i];
    }
    lam = den > 0 ? num / den : 1.0;
    nrm = std::sqrt(nrm);
    if (!(nrm > 0)) break;
    for (int i = 0; i < n; ++i) x[i] = y[i] / nrm;
  }
  return lam > 0 ? lam : 1.0;
}

// Three-phase greedy aggregation over the strong-connection graph.
std::vector<int> aggregate(const HostCsr& A, const std::vector<double>& d, double theta, int* n_agg) {
  const int n = A.n;
  std::vector<int> sptr((size_t)n + 1, 0), scol;
  std::vector<double> sval;
  for (int i = 0; i < n; ++i) {
    for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
      const int j = A.col[k];
      if (j == i) continue;
      const double a = std::fabs(A.val[k]);
      if (a >= theta * std::sqrt(std::fabs(d[i] * d[j])) && a > 0) {
        scol.push_back(j);
        sval.push_back(a);
      }
    }
    sptr[i + 1] = (int)scol.size();
  }
  std::vector<int> agg((size_t)n, -1);
  int na = 0;
  // phase 1: whole neighbourhoods that are still free
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    bool free_nb = true;
    for (int k = sptr[i]; k < sptr[i + 1] && free_nb; ++k) free_nb = agg[scol[k]] < 0;
    if (!free_nb) continue;
    agg[i] = na;
    for (int k = sptr[i]; k < sptr[i + 1]; ++k) agg[scol[k]] = na;
    ++na;
  }
  // phase 2: join the aggregate of the strongest already aggregated neighbour
  std::vector<int> agg1 = agg;
  for (int i = 0; i < n; ++i) {
    if (agg1[i] >= 0) continue;
    double best = -1;
    int to = -1;
    for (int k = sptr[i]; k < sptr[i + 1]; ++k)
      if (agg1[scol[k]] >= 0 && sval[k] > best) {
        best = sval[k];
        to = agg1[scol[k]];
      }
    agg[i] = to;
  }
  // phase 3: what is left forms aggregates with its free strong neighbours
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    agg[i] = na;
    for (int k = sptr[i]; k < sptr[i + 1]; ++k)
      if (agg[scol[k]] < 0) agg[scol[k]] = na;
    ++na;
  }
  *n_agg = na;
  return agg;
}

// Cholesky factorisation and explicit inverse of a dense SPD matrix (row-major).
bool dense_spd_inverse(std::vector<double> a, int n, std::vector<double>& inv) {
  for (int j = 0; j < n; ++j) {
    double s = a[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) s -= a[(size_t)j * n + k] * a[(size_t)j * n + k];
    if (!(s > 0)) return false;
    const double ljj = std::sqrt(s);
    a[(size_t)j * n + j] = ljj;
    for (int i = j + 1; i < n; ++i) {
      double t = a[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) t -= a[(size_t)i * n + k] * a[(size_t)j * n + k];
      a[(size_t)i * n + j] = t / ljj;
    }
  }
  // inverse column by column: L L^T x = e_c
  inv.assign((size_t)n * n, 0.0);
  std::vector<double> y((size_t)n);
  for (int c = 0; c < n; ++c) {
    for (int i = 0; i < n; ++i) {
      double t = i == c ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) t -= a[(size_t)i * n + k] * y[k];
      y[i] = t / a[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double t = y[i];
      for (int k = i + 1; k < n; ++k) t -= a[(size_t)k * n + i] * inv[(size_t)k * n + c];
      inv[(size_t)i * n + c] = t / a[(size_t)i * n + i];
    }
  }
  return true;
}

}  // namespace

AmgHierarchy amg_setup(HostCsr A, const AmgOptions& opt, bool singular) {
  AmgHierarchy H;
  for (int lev = 0;; ++lev) {
    AmgHostLevel L;
    L.diag = diagonal(A);
    L.lmax = lambda_max(A, L.diag, 15);
    const bool last = A.n <= opt.max_coarse || lev + 1 >= opt.max_levels;
    if (!last) {
      int na = 0;
      std::vector<int> agg = aggregate(A, L.diag, opt.theta, &na);
      if (na > 0.8 * A.n) agg = aggregate(A, L.diag, 0.0, &na);  // stalled: every entry is strong
      if (na >= A.n || na == 0) {
        L.A = std::move(A);
        H.levels.push_back(std::move(L));
        break;
      }
      // tentative prolongator: orthonormal indicator columns of the aggregates
      std::vector<int> size((size_t)na, 0);
      for (int i = 0; i < A.n; ++i) ++size[agg[i]];
      HostCsr Pt;
      Pt.n = A.n;
      Pt.m = na;
      Pt.ptr.resize((size_t)A.n + 1);
      Pt.col.resize((size_t)A.n);
      Pt.val.resize((size_t)A.n);
      for (int i = 0; i < A.n; ++i) {
        Pt.ptr[i] = i;
        Pt.col[i] = agg[i];
        Pt.val[i] = 1.0 / std::sqrt((double)size[agg[i]]);
      }
      Pt.ptr[A.n] = A.n;
      // P = (I - omega/lambda D^-1 A) Pt
      HostCsr AP = csr_multiply(A, Pt);
      const double w = opt.omega / L.lmax;
      HostCsr P;
      P.n = A.n;
      P.m = na;
      P.ptr.assign((size_t)A.n + 1, 0);
      for (int i = 0; i < A.n; ++i) {
        const double s = L.diag[i] != 0 ? w / L.diag[i] : 0.0;
        // merge row i of Pt (one entry) with -s * row i of AP (sorted columns)
        bool placed = false;
        for (int k = AP.ptr[i]; k < AP.ptr[i + 1]; ++k) {
          const int c = AP.col[k];
          double v = -s * AP.val[k];
          if (!placed && c >= Pt.col[i]) {
            if (c == Pt.col[i]) {
              v += Pt.val[i];
            } else {
              P.col.push_back(Pt.col[i]);
              P.val.push_back(Pt.val[i]);
            }
            placed = true;
          }
          P.col.push_back(c);
          P.val.push_back(v);
        }
        if (!placed) {
          P.col.push_back(Pt.col[i]);
          P.val.push_back(Pt.val[i]);
        }
        P.ptr[i + 1] = (int)P.col.size();
      }
      HostCsr R = csr_transpose(P);
      HostCsr Ac = csr_multiply(R, csr_multiply(A, P));
      // symmetrise against rounding
      HostCsr AcT = csr_transpose(Ac);
      if (AcT.ptr == Ac.ptr && AcT.col == Ac.col)
        for (size_t k = 0; k < Ac.val.size(); ++k) Ac.val[k] = 0.5 * (Ac.val[k] + AcT.val[k]);
      L.A = std::move(A);
      L.P = std::move(P);
      L.R = std::move(R);
      H.levels.push_back(std::move(L));
      A = std::move(Ac);
      continue;
    }
    L.A = std::move(A);
    H.levels.push_back(std::move(L));
    break;
  }
  // dense inverse on the coarsest level
  const HostCsr& C = H.levels.back().A;
  const int n = C.n;
  if (n > 4096) throw std::runtime_error("AMG coarsening stalled above the dense-solve size");
  std::vector<double> dense((size_t)n * n, 0.0);
  double tr = 0;
  for (int i = 0; i < n; ++i)
    for (int k = C.ptr[i]; k < C.ptr[i + 1]; ++k) {
      dense[(size_t)i * n + C.col[k]] = C.val[k];
      if (C.col[k] == i) tr += C.val[k];
    }
  const double mean_diag = n ? tr / n : 1.0;
  if (singular)
    for (auto& v : dense) v += mean_diag / n;
  std::vector<double> inv;
  bool ok = dense_spd_inverse(dense, n, inv);
  for (double shift = 1e-12; !ok && shift < 1e-2; shift *= 100) {
    std::vector<double> t = dense;
    for (int i = 0; i < n; ++i) t[(size_t)i * n + i] += shift * mean_diag;
    ok = dense_spd_inverse(t, n, inv);
  }
  if (!ok) throw std::runtime_error("coarsest AMG operator is not positive definite");
  H.n_coarse = n;
  H.coarse_inv = std::move(inv);
  return H;
}

}  // namespace insp
