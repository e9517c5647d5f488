/*
 * FP64 CPU oracle for the DG Poisson solve - TEST INFRASTRUCTURE ONLY.
 * See ins_oracle.h. Every operator step cites the reference kernel it restates.
 *
 * Pinning of each function (checked in tests/):
 *  - nodes: N=1 vertices, 2D N=2 equidistant, edge nodes = Gauss-Lobatto points,
 *    symmetry under vertex permutations; cross-checked against the product's
 *    independently coded construction;
 *  - D, M, face mass: exactness on monomials (closed forms), M.1 = |T|,
 *    face mass sums = |face|;
 *  - mult: symmetry, SPD (Dirichlet), A.1 = 0 (Neumann), O(h^{N+1})
 *    manufactured-solution convergence, and the dense operator equal (1e-12) to an
 *    independent SIPG bilinear form assembled by quadrature with its own basis,
 *    face search and penalty (tests/sipg_textbook.py, tests/test_cpu_oracle_pins.py:
 *    2D/3D, N=1..4, Dirichlet/Neumann/mixed, constant/variable k; the same
 *    assembly with tau or the symmetry term off by 2x is shown to differ);
 *  - solvers: dense Cholesky solve on tiny meshes; p-MG converges to the same
 *    solution as Jacobi-PCG; the power iteration converges to the top eigenvalue
 *    of D^-1 A from a dense eigen-solve and its 15-step estimate lies below it.
 *  - mass term: A.1 - A_0.1 = J M mm exactly (Neumann, where the SIPG part of A.1
 *    vanishes), its diagonal is the diagonal of the assembled term, Helmholtz
 *    manufactured solutions converge at O(h^{N+1}); inverse-mass PCG reaches the
 *    dense solution.
 */
#define _GNU_SOURCE
#include "ins_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXNP 35
#define MAXNFP 15

/* ------------------------------------------------------------------ */
/* dense helpers                                                       */
/* ------------------------------------------------------------------ */

static int inv_dense(const double* A, int n, double* out) {
  double* a = malloc(sizeof(double) * n * n);
  memcpy(a, A, sizeof(double) * n * n);
  for (int i = 0; i < n * n; ++i) out[i] = 0;
  for (int i = 0; i < n; ++i) out[i * n + i] = 1;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int i = c + 1; i < n; ++i)
      if (fabs(a[i * n + c]) > fabs(a[piv * n + c])) piv = i;
    if (fabs(a[piv * n + c]) < 1e-300) { free(a); return -1; }
    if (piv != c)
      for (int j = 0; j < n; ++j) {
        double t = a[c * n + j]; a[c * n + j] = a[piv * n + j]; a[piv * n + j] = t;
        t = out[c * n + j]; out[c * n + j] = out[piv * n + j]; out[piv * n + j] = t;
      }
    const double d = 1.0 / a[c * n + c];
    for (int j = 0; j < n; ++j) { a[c * n + j] *= d; out[c * n + j] *= d; }
    for (int i = 0; i < n; ++i) {
      if (i == c) continue;
      const double f = a[i * n + c];
      if (f == 0) continue;
      for (int j = 0; j < n; ++j) { a[i * n + j] -= f * a[c * n + j]; out[i * n + j] -= f * out[c * n + j]; }
    }
  }
  free(a);
  return 0;
}

static double fact(int n) {
  double f = 1;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}

/* ------------------------------------------------------------------ */
/* nodes (Hesthaven & Warburton 2008, Nodes2D / Nodes3D)               */
/* ------------------------------------------------------------------ */

/* Legendre-Gauss-Lobatto points, ascending (Newton on (1-x^2)P_N'). */
static void gll(int N, double* x) {
  for (int k = 0; k <= N; ++k) {
    double z = -cos(M_PI * k / N), zold;
    int it = 0;
    do {
      zold = z;
      double p0 = 1, p1 = z;
      for (int n = 2; n <= N; ++n) {
        const double p2 = ((2 * n - 1) * z * p1 - (n - 1) * p0) / n;
        p0 = p1; p1 = p2;
      }
      if (N == 1) { p0 = 1; p1 = z; }
      z = zold - (z * p1 - p0) / ((N + 1) * p1);
    } while (fabs(z - zold) > 1e-16 && ++it < 100);
    x[k] = z;
  }
  x[0] = -1; x[N] = 1;
}

static double legendre(int n, double x) {
  double p0 = 1, p1 = x;
  if (n == 0) return 1;
  for (int k = 2; k <= n; ++k) { const double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
  return p1;
}

/* Warp factor: Legendre interpolant of (GLL - equidistant) divided by 1 - r^2. */
static double warpfactor(int N, double rout) {
  double lgl[16], req[16], V[256] = {0}, Vi[256], c[16];
  gll(N, lgl);
  for (int i = 0; i <= N; ++i) req[i] = -1 + 2.0 * i / N;
  for (int i = 0; i <= N; ++i)
    for (int j = 0; j <= N; ++j) V[i * (N + 1) + j] = legendre(j, req[i]);
  inv_dense(V, N + 1, Vi);
  for (int j = 0; j <= N; ++j) {
    c[j] = 0;
    for (int i = 0; i <= N; ++i) c[j] += Vi[j * (N + 1) + i] * (lgl[i] - req[i]);
  }
  double w = 0;
  for (int j = 0; j <= N; ++j) w += c[j] * legendre(j, rout);
  if (fabs(rout) < 1.0 - 1e-10) return w / (1 - rout * rout);
  return 0;
}

static void nodes2d(int N, double* r, double* s) {
  static const double alpopt[15] = {0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832,
                                    1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258};
  const double alpha = alpopt[N - 1];
  int sk = 0;
  for (int n = 0; n <= N; ++n)
    for (int m = 0; m <= N - n; ++m) {
      const double L1 = (double)n / N, L3 = (double)m / N, L2 = 1 - L1 - L3;
      double x = -L2 + L3, y = (-L2 - L3 + 2 * L1) / sqrt(3.0);
      const double w1 = 4 * L2 * L3 * warpfactor(N, L3 - L2) * (1 + (alpha * L1) * (alpha * L1));
      const double w2 = 4 * L1 * L3 * warpfactor(N, L1 - L3) * (1 + (alpha * L2) * (alpha * L2));
      const double w3 = 4 * L1 * L2 * warpfactor(N, L2 - L1) * (1 + (alpha * L3) * (alpha * L3));
      x += w1 + cos(2 * M_PI / 3) * w2 + cos(4 * M_PI / 3) * w3;
      y += sin(2 * M_PI / 3) * w2 + sin(4 * M_PI / 3) * w3;
      const double l1 = (sqrt(3.0) * y + 1) / 3, l2 = (-3 * x - sqrt(3.0) * y + 2) / 6,
                   l3 = (3 * x - sqrt(3.0) * y + 2) / 6;
      r[sk] = -l2 + l3 - l1;
      s[sk] = -l2 - l3 + l1;
      ++sk;
    }
}

static void shift3(int N, double pval, double L1, double L2, double L3, double* dx, double* dy) {
  const double f1 = 4 * L2 * L3 * warpfactor(N, L3 - L2) * (1 + (pval * L1) * (pval * L1));
  const double f2 = 4 * L1 * L3 * warpfactor(N, L1 - L3) * (1 + (pval * L2) * (pval * L2));
  const double f3 = 4 * L1 * L2 * warpfactor(N, L2 - L1) * (1 + (pval * L3) * (pval * L3));
  *dx = f1 + cos(2 * M_PI / 3) * f2 + cos(4 * M_PI / 3) * f3;
  *dy = sin(2 * M_PI / 3) * f2 + sin(4 * M_PI / 3) * f3;
}

static void nodes3d(int N, double* r, double* s, double* t) {
  static const double alpopt[15] = {0, 0, 0, 0.1002, 1.1332, 1.5608, 1.3413, 1.2577,
                                    1.1603, 1.10153, 0.6080, 0.4523, 0.8856, 0.8717, 0.9655};
  const double alpha = alpopt[N - 1], tol = 1e-10;
  const double v[4][3] = {{-1, -1 / sqrt(3.0), -1 / sqrt(6.0)}, {1, -1 / sqrt(3.0), -1 / sqrt(6.0)},
                          {0, 2 / sqrt(3.0), -1 / sqrt(6.0)}, {0, 0, 3 / sqrt(6.0)}};
  double t1[4][3], t2[4][3];
  for (int k = 0; k < 3; ++k) {
    t1[0][k] = v[1][k] - v[0][k]; t1[1][k] = v[1][k] - v[0][k];
    t1[2][k] = v[2][k] - v[1][k]; t1[3][k] = v[2][k] - v[0][k];
    t2[0][k] = v[2][k] - 0.5 * (v[0][k] + v[1][k]); t2[1][k] = v[3][k] - 0.5 * (v[0][k] + v[1][k]);
    t2[2][k] = v[3][k] - 0.5 * (v[1][k] + v[2][k]); t2[3][k] = v[3][k] - 0.5 * (v[0][k] + v[2][k]);
  }
  for (int f = 0; f < 4; ++f) {
    double a = 0, b = 0;
    for (int k = 0; k < 3; ++k) { a += t1[f][k] * t1[f][k]; b += t2[f][k] * t2[f][k]; }
    for (int k = 0; k < 3; ++k) { t1[f][k] /= sqrt(a); t2[f][k] /= sqrt(b); }
  }
  /* A maps (r,s,t)+1 to equilateral coordinates: X = v0 + sum_k (v_k - v0)(x_k+1)/2 */
  double A[9], Ai[9];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) A[i * 3 + k] = 0.5 * (v[k + 1][i] - v[0][i]);
  inv_dense(A, 3, Ai);
  int sk = 0;
  for (int n = 0; n <= N; ++n)
    for (int m = 0; m <= N - n; ++m)
      for (int q = 0; q <= N - n - m; ++q) {
        const double rr = -1 + 2.0 * q / N, ss = -1 + 2.0 * m / N, tt = -1 + 2.0 * n / N;
        const double L1 = (1 + tt) / 2, L2 = (1 + ss) / 2, L3 = -(1 + rr + ss + tt) / 2, L4 = (1 + rr) / 2;
        double X[3], sh[3] = {0, 0, 0};
        for (int k = 0; k < 3; ++k) X[k] = L3 * v[0][k] + L4 * v[1][k] + L2 * v[2][k] + L1 * v[3][k];
        for (int f = 0; f < 4; ++f) {
          double La, Lb, Lc, Ld;
          if (f == 0) { La = L1; Lb = L2; Lc = L3; Ld = L4; }
          else if (f == 1) { La = L2; Lb = L1; Lc = L3; Ld = L4; }
          else if (f == 2) { La = L3; Lb = L1; Lc = L4; Ld = L2; }
          else { La = L4; Lb = L1; Lc = L3; Ld = L2; }
          double w1, w2;
          shift3(N, alpha, Lb, Lc, Ld, &w1, &w2);
          double blend = Lb * Lc * Ld;
          const double den = (Lb + 0.5 * La) * (Lc + 0.5 * La) * (Ld + 0.5 * La);
          if (den > tol) blend = (1 + (alpha * La) * (alpha * La)) * blend / den;
          if (La < tol && ((Lb > tol) + (Lc > tol) + (Ld > tol)) < 3)
            for (int k = 0; k < 3; ++k) sh[k] = w1 * t1[f][k] + w2 * t2[f][k];
          else
            for (int k = 0; k < 3; ++k) sh[k] += blend * (w1 * t1[f][k] + w2 * t2[f][k]);
        }
        double rel[3];
        for (int k = 0; k < 3; ++k) rel[k] = X[k] + sh[k] - v[0][k];
        double rst[3];
        for (int i = 0; i < 3; ++i) rst[i] = Ai[i * 3] * rel[0] + Ai[i * 3 + 1] * rel[1] + Ai[i * 3 + 2] * rel[2] - 1;
        r[sk] = rst[0]; s[sk] = rst[1]; t[sk] = rst[2];
        ++sk;
      }
}

/* ------------------------------------------------------------------ */
/* reference element from monomials on the unit simplex               */
/* ------------------------------------------------------------------ */

typedef struct {
  int dim, N, Np, Nfp, Nf;
  double r[MAXNP], s[MAXNP], t[MAXNP];
  double D[3][MAXNP * MAXNP];
  double M[MAXNP * MAXNP];
  int fmask[4 * MAXNFP];
  double Mface[4 * MAXNFP * MAXNFP];
  double MF[4][MAXNP * MAXNP]; /* face mass as Np x Np (zero off the face) */
  double EMAT[MAXNP * 4 * MAXNFP];
  double Vi[MAXNP * MAXNP];  /* inverse monomial Vandermonde: phi_j = sum_k m_k Vi[k][j] */
  int ex[MAXNP][3];          /* monomial exponents in x = (1 + r) / 2 */
} refel;

static const int FV2[3][2] = {{0, 1}, {1, 2}, {2, 0}};
static const int FV3[4][3] = {{0, 1, 2}, {0, 1, 3}, {1, 2, 3}, {0, 2, 3}};

static int monomials(int dim, int N, int (*ex)[3]) {
  int n = 0;
  for (int a = 0; a <= N; ++a)
    for (int b = 0; b <= N - a; ++b)
      if (dim == 2) { ex[n][0] = a; ex[n][1] = b; ex[n][2] = 0; ++n; }
      else for (int c = 0; c <= N - a - b; ++c) { ex[n][0] = a; ex[n][1] = b; ex[n][2] = c; ++n; }
  return n;
}

/* integral of x^a y^b z^c over the unit simplex of dimension dim */
static double simplex_int(int dim, int a, int b, int c) {
  return fact(a) * fact(b) * fact(c) / fact(a + b + c + dim);
}

static void barycentric(int dim, double r, double s, double t, double* lam) {
  if (dim == 2) { lam[0] = -(r + s) / 2; lam[1] = (1 + r) / 2; lam[2] = (1 + s) / 2; }
  else { lam[0] = -(1 + r + s + t) / 2; lam[1] = (1 + r) / 2; lam[2] = (1 + s) / 2; lam[3] = (1 + t) / 2; }
}

static int build_refel(int dim, int N, refel* R) {
  memset(R, 0, sizeof(*R));
  R->dim = dim; R->N = N; R->Nf = dim + 1;
  R->Np = dim == 2 ? (N + 1) * (N + 2) / 2 : (N + 1) * (N + 2) * (N + 3) / 6;
  R->Nfp = dim == 2 ? N + 1 : (N + 1) * (N + 2) / 2;
  const int Np = R->Np, Nfp = R->Nfp, Nf = R->Nf;
  if (dim == 2) nodes2d(N, R->r, R->s); else nodes3d(N, R->r, R->s, R->t);
  int ex[MAXNP][3];
  monomials(dim, N, ex);
  double V[MAXNP * MAXNP], Vi[MAXNP * MAXNP], Vd[MAXNP * MAXNP];
  for (int i = 0; i < Np; ++i) {
    const double x[3] = {(1 + R->r[i]) / 2, (1 + R->s[i]) / 2, (1 + R->t[i]) / 2};
    for (int k = 0; k < Np; ++k)
      V[i * Np + k] = pow(x[0], ex[k][0]) * pow(x[1], ex[k][1]) * (dim == 3 ? pow(x[2], ex[k][2]) : 1.0);
  }
  if (inv_dense(V, Np, Vi)) return -1;
  memcpy(R->Vi, Vi, sizeof(double) * Np * Np);
  memcpy(R->ex, ex, sizeof(int) * 3 * Np);
  for (int d = 0; d < dim; ++d) {
    for (int i = 0; i < Np; ++i) {
      const double x[3] = {(1 + R->r[i]) / 2, (1 + R->s[i]) / 2, (1 + R->t[i]) / 2};
      for (int k = 0; k < Np; ++k) {
        int e[3] = {ex[k][0], ex[k][1], ex[k][2]};
        double v = 0;
        if (e[d] > 0) {
          v = 0.5 * e[d];  /* d/dr = 1/2 d/dx */
          e[d] -= 1;
          v *= pow(x[0], e[0]) * pow(x[1], e[1]) * (dim == 3 ? pow(x[2], e[2]) : 1.0);
        }
        Vd[i * Np + k] = v;
      }
    }
    for (int i = 0; i < Np; ++i)
      for (int j = 0; j < Np; ++j) {
        double a = 0;
        for (int k = 0; k < Np; ++k) a += Vd[i * Np + k] * Vi[k * Np + j];
        R->D[d][i * Np + j] = a;
      }
  }
  /* M_ij = sum_kl Vi[k][i] Vi[l][j] int(m_k m_l) * 2^dim */
  const double jac = dim == 2 ? 4.0 : 8.0;
  double G[MAXNP * MAXNP];
  for (int k = 0; k < Np; ++k)
    for (int l = 0; l < Np; ++l)
      G[k * Np + l] = jac * simplex_int(dim, ex[k][0] + ex[l][0], ex[k][1] + ex[l][1], ex[k][2] + ex[l][2]);
  for (int i = 0; i < Np; ++i)
    for (int j = 0; j < Np; ++j) {
      double a = 0;
      for (int k = 0; k < Np; ++k)
        for (int l = 0; l < Np; ++l) a += Vi[k * Np + i] * G[k * Np + l] * Vi[l * Np + j];
      R->M[i * Np + j] = a;
    }
  /* faces */
  for (int f = 0; f < Nf; ++f) {
    int opp = -1;
    for (int vtx = 0; vtx <= dim; ++vtx) {
      int on = 0;
      for (int k = 0; k < dim; ++k) on |= (dim == 2 ? FV2[f][k] : FV3[f][k]) == vtx;
      if (!on) opp = vtx;
    }
    int cnt = 0;
    double fb[MAXNFP][2];
    for (int i = 0; i < Np; ++i) {
      double lam[4];
      barycentric(dim, R->r[i], R->s[i], R->t[i], lam);
      if (fabs(lam[opp]) < 1e-10) {
        if (cnt >= Nfp) return -2;
        R->fmask[f * Nfp + cnt] = i;
        /* face-local coordinates: barycentrics of the 2nd (and 3rd) face vertex */
        fb[cnt][0] = lam[dim == 2 ? FV2[f][1] : FV3[f][1]];
        fb[cnt][1] = dim == 3 ? lam[FV3[f][2]] : 0;
        ++cnt;
      }
    }
    if (cnt != Nfp) return -3;
    int fex[MAXNFP][3];
    monomials(dim - 1 == 1 ? 2 : 2, N, fex); /* 2D monomials; 1D uses those with b = 0 */
    int nfm = 0, fe[MAXNFP][2];
    for (int k = 0; k < (N + 1) * (N + 2) / 2; ++k) {
      if (dim == 2 && fex[k][1] != 0) continue;
      fe[nfm][0] = fex[k][0]; fe[nfm][1] = fex[k][1]; ++nfm;
    }
    double Vf[MAXNFP * MAXNFP], Vfi[MAXNFP * MAXNFP];
    for (int j = 0; j < Nfp; ++j)
      for (int k = 0; k < Nfp; ++k) Vf[j * Nfp + k] = pow(fb[j][0], fe[k][0]) * pow(fb[j][1], fe[k][1]);
    if (inv_dense(Vf, Nfp, Vfi)) return -4;
    const double fjac = dim == 2 ? 2.0 : 4.0; /* reference face measure / unit simplex measure */
    for (int i = 0; i < Nfp; ++i)
      for (int j = 0; j < Nfp; ++j) {
        double a = 0;
        for (int k = 0; k < Nfp; ++k)
          for (int l = 0; l < Nfp; ++l)
            a += Vfi[k * Nfp + i] * Vfi[l * Nfp + j] * fjac *
                 simplex_int(dim - 1, fe[k][0] + fe[l][0], fe[k][1] + fe[l][1], 0);
        R->Mface[(f * Nfp + i) * Nfp + j] = a;
        R->MF[f][R->fmask[f * Nfp + i] * Np + R->fmask[f * Nfp + j]] = a;
        R->EMAT[R->fmask[f * Nfp + i] * (Nf * Nfp) + f * Nfp + j] = a;
      }
  }
  return 0;
}

static void interp_mono(int dim, const refel* F, const refel* Cc, double* P) {
  int ex[MAXNP][3];
  const int npc = Cc->Np, npf = F->Np;
  monomials(dim, Cc->N, ex);
  double V[MAXNP * MAXNP], Vi[MAXNP * MAXNP];
  for (int i = 0; i < npc; ++i) {
    const double x[3] = {(1 + Cc->r[i]) / 2, (1 + Cc->s[i]) / 2, (1 + Cc->t[i]) / 2};
    for (int k = 0; k < npc; ++k)
      V[i * npc + k] = pow(x[0], ex[k][0]) * pow(x[1], ex[k][1]) * (dim == 3 ? pow(x[2], ex[k][2]) : 1.0);
  }
  inv_dense(V, npc, Vi);
  for (int i = 0; i < npf; ++i) {
    const double x[3] = {(1 + F->r[i]) / 2, (1 + F->s[i]) / 2, (1 + F->t[i]) / 2};
    for (int j = 0; j < npc; ++j) {
      double a = 0;
      for (int k = 0; k < npc; ++k)
        a += pow(x[0], ex[k][0]) * pow(x[1], ex[k][1]) * (dim == 3 ? pow(x[2], ex[k][2]) : 1.0) * Vi[k * npc + j];
      P[i * npc + j] = a;
    }
  }
}

int oracle_ref_element(int dim, int order, double* rst, double* D, double* M, int* fmask,
                       double* Mface) {
  refel* R = malloc(sizeof(refel));
  if (build_refel(dim, order, R)) { free(R); return -1; }
  const int Np = R->Np;
  if (rst)
    for (int i = 0; i < Np; ++i) {
      rst[i] = R->r[i]; rst[Np + i] = R->s[i];
      if (dim == 3) rst[2 * Np + i] = R->t[i];
    }
  if (D)
    for (int d = 0; d < dim; ++d) memcpy(D + (size_t)d * Np * Np, R->D[d], sizeof(double) * Np * Np);
  if (M) memcpy(M, R->M, sizeof(double) * Np * Np);
  if (fmask) memcpy(fmask, R->fmask, sizeof(int) * R->Nf * R->Nfp);
  if (Mface) memcpy(Mface, R->Mface, sizeof(double) * R->Nf * R->Nfp * R->Nfp);
  free(R);
  return 0;
}

int oracle_interp(int dim, int nf, int nc, double* P) {
  refel* F = malloc(sizeof(refel));
  refel* Cc = malloc(sizeof(refel));
  int rc = build_refel(dim, nf, F) || build_refel(dim, nc, Cc);
  if (!rc) interp_mono(dim, F, Cc, P);
  free(F); free(Cc);
  return rc ? -1 : 0;
}

/* ------------------------------------------------------------------ */
/* mesh, faces and geometry                                            */
/* ------------------------------------------------------------------ */

typedef struct {
  int N;
  refel R;
  double* tau;   /* E x Nf  (mat_free_tau_c) */
  double* k;     /* E x Np or NULL */
  double* diag;  /* E x Np */
  double* inv;   /* E x Np */
  double lmax;
  double* P;     /* to next coarser: Npf x Npc */
  int* map;      /* face node matching at this order */
} level_t;

struct oracle {
  int dim, N;
  int64_t nn, E, F, B;
  double* coords;
  int* c2n;
  refel R;
  double* geof;  /* E x (dim*dim + 1): G[r][x] = d xi_r / d x_x, then J */
  double* nrm;   /* E x Nf x dim */
  double* sJ;    /* E x Nf */
  double* fsc;   /* E x Nf */
  int *f2c, *fnum, *b2c, *bnum, *bct;
  int* mapLpos;  /* F x Nfp: position (in R's face) matching L face node j, element order */
  int* cfb;      /* E x Nf: bface index of a cell face, or -1 */
  double* k;
  int pen_order; /* order in the penalty 2(N+1)(N+2) */
  double* mm;    /* E x Np mass-term factor (FactorMMPoissonMatrixFree) or NULL */
  double mmc;    /* constant mass-term factor (MMPoissonMatrixFree), 0 = none */
};

static int g_threads = 0;
void oracle_set_threads(int n) { g_threads = n; }
#define OMP_FOR _Pragma("omp parallel for schedule(static) if(g_threads != 1) num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())")
#ifndef _OPENMP
static int omp_get_max_threads(void) { return 1; }
#endif

typedef struct { int key[3]; int64_t cell; int face; } facerec;

static int cmp_face(const void* a, const void* b) {
  const facerec *x = a, *y = b;
  for (int i = 0; i < 3; ++i)
    if (x->key[i] != y->key[i]) return x->key[i] < y->key[i] ? -1 : 1;
  if (x->cell != y->cell) return x->cell < y->cell ? -1 : 1;
  return x->face - y->face;
}

static int cmp_pair(const void* a, const void* b) {
  const int64_t *x = a, *y = b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  return (int)(x[1] - y[1]);
}

static void vsub(const double* a, const double* b, double* o, int d) { for (int i = 0; i < d; ++i) o[i] = a[i] - b[i]; }

/* For each interior face, the position in R's face mask of the node that coincides with
 * face node j of L (matched by physical coordinates, as the toolkit's fmaskR does; relative
 * to each side's face centroid, so a periodic face, whose sides differ by a translation,
 * matches as well). */
static int* face_map(const oracle_t* o, const refel* R) {
  const int dim = o->dim, nv = dim + 1, Nfp = R->Nfp;
  int* map = malloc(sizeof(int) * Nfp * (o->F + 1));
  for (int64_t i = 0; i < o->F; ++i) {
    const int L = o->f2c[2 * i], Rc = o->f2c[2 * i + 1], fL = o->fnum[2 * i], fR = o->fnum[2 * i + 1];
    const int* cL = o->c2n + (size_t)L * nv;
    const int* cR = o->c2n + (size_t)Rc * nv;
    const int* fvL = dim == 2 ? FV2[fL] : FV3[fL];
    const int* fvR = dim == 2 ? FV2[fR] : FV3[fR];
    double ctrL[3] = {0, 0, 0}, ctrR[3] = {0, 0, 0};
    for (int k = 0; k < dim; ++k)
      for (int x = 0; x < dim; ++x) {
        ctrL[x] += o->coords[(size_t)cL[fvL[k]] * dim + x] / dim;
        ctrR[x] += o->coords[(size_t)cR[fvR[k]] * dim + x] / dim;
      }
    double scale = 0;
    for (int x = 0; x < dim; ++x) {
      const double d = o->coords[(size_t)cL[1] * dim + x] - o->coords[(size_t)cL[0] * dim + x];
      scale += d * d;
    }
    scale = sqrt(scale);
    for (int j = 0; j < Nfp; ++j) {
      const int iL = R->fmask[fL * Nfp + j];
      double lamL[4], xL[3] = {0, 0, 0};
      barycentric(dim, R->r[iL], R->s[iL], R->t[iL], lamL);
      for (int v = 0; v < nv; ++v) for (int x = 0; x < dim; ++x) xL[x] += lamL[v] * o->coords[(size_t)cL[v] * dim + x];
      int found = -1;
      for (int q = 0; q < Nfp; ++q) {
        const int iR = R->fmask[fR * Nfp + q];
        double lamR[4], xR[3] = {0, 0, 0};
        barycentric(dim, R->r[iR], R->s[iR], R->t[iR], lamR);
        for (int v = 0; v < nv; ++v) for (int x = 0; x < dim; ++x) xR[x] += lamR[v] * o->coords[(size_t)cR[v] * dim + x];
        double d2 = 0;
        for (int x = 0; x < dim; ++x) {
          const double dx = (xL[x] - ctrL[x]) - (xR[x] - ctrR[x]);
          d2 += dx * dx;
        }
        if (sqrt(d2) < 1e-9 * scale) found = q;
      }
      if (found < 0) { free(map); return NULL; }
      map[i * Nfp + j] = found;
    }
  }
  return map;
}

static oracle_t* create(int dim, int order, int64_t n_nodes, const double* coords, int64_t n_cells,
                        const int* cell2nodes, int64_t cF, const int* cf2c, const int* cfnum, int64_t cB,
                        const int* cb2c, const int* cbnum);

oracle_t* oracle_create(int dim, int order, int64_t n_nodes, const double* coords, int64_t n_cells,
                        const int* cell2nodes) {
  return create(dim, order, n_nodes, coords, n_cells, cell2nodes, -1, NULL, NULL, -1, NULL, NULL);
}

/* The caller's face sets (the OP2 mesh file's face2cells/faceNum and bface2cells/bfaceNum,
 * src/ins2d.cpp:192) in place of the oracle's own face search: faces may join cells whose
 * face vertices differ by a translation (periodic faces). Cells must be positively oriented. */
oracle_t* oracle_create_faces(int dim, int order, int64_t n_nodes, const double* coords, int64_t n_cells,
                              const int* cell2nodes, int64_t n_faces, const int* face2cells,
                              const int* faceNum, int64_t n_bfaces, const int* bface2cells,
                              const int* bfaceNum) {
  if (n_faces < 0 || n_bfaces < 0) return NULL;
  return create(dim, order, n_nodes, coords, n_cells, cell2nodes, n_faces, face2cells, faceNum, n_bfaces,
                bface2cells, bfaceNum);
}

static oracle_t* create(int dim, int order, int64_t n_nodes, const double* coords, int64_t n_cells,
                        const int* cell2nodes, int64_t cF, const int* cf2c, const int* cfnum, int64_t cB,
                        const int* cb2c, const int* cbnum) {
  if ((dim != 2 && dim != 3) || order < 1 || order > 4) return NULL;
  oracle_t* o = calloc(1, sizeof(oracle_t));
  o->dim = dim; o->N = order; o->nn = n_nodes; o->E = n_cells; o->pen_order = order;
  const int nv = dim + 1, Nf = dim + 1;
  o->coords = malloc(sizeof(double) * n_nodes * dim);
  memcpy(o->coords, coords, sizeof(double) * n_nodes * dim);
  o->c2n = malloc(sizeof(int) * n_cells * nv);
  memcpy(o->c2n, cell2nodes, sizeof(int) * n_cells * nv);
  if (build_refel(dim, order, &o->R)) { oracle_destroy(o); return NULL; }
  const int64_t E = n_cells;
  /* orientation: positive measure (swap the last two vertices otherwise) */
  for (int64_t e = 0; e < E; ++e) {
    int* c = o->c2n + e * nv;
    double a[3] = {0, 0, 0}, b[3] = {0, 0, 0}, cc[3] = {0, 0, 0}, vol;
    vsub(o->coords + (size_t)c[1] * dim, o->coords + (size_t)c[0] * dim, a, dim);
    vsub(o->coords + (size_t)c[2] * dim, o->coords + (size_t)c[0] * dim, b, dim);
    if (dim == 2) vol = a[0] * b[1] - a[1] * b[0];
    else {
      vsub(o->coords + (size_t)c[3] * dim, o->coords + (size_t)c[0] * dim, cc, dim);
      vol = a[0] * (b[1] * cc[2] - b[2] * cc[1]) - a[1] * (b[0] * cc[2] - b[2] * cc[0]) + a[2] * (b[0] * cc[1] - b[1] * cc[0]);
    }
    if (vol < 0) {
      if (cF >= 0) { oracle_destroy(o); return NULL; }  /* caller's faces need its vertex order */
      int t = c[dim - 1]; c[dim - 1] = c[dim]; c[dim] = t;
    }
  }
  /* geometry */
  const int ng = dim * dim + 1;
  o->geof = malloc(sizeof(double) * E * ng);
  o->nrm = malloc(sizeof(double) * E * Nf * dim);
  o->sJ = malloc(sizeof(double) * E * Nf);
  o->fsc = malloc(sizeof(double) * E * Nf);
  for (int64_t e = 0; e < E; ++e) {
    const int* c = o->c2n + e * nv;
    double X[4][3];
    for (int v = 0; v < nv; ++v)
      for (int x = 0; x < dim; ++x) X[v][x] = o->coords[(size_t)c[v] * dim + x];
    /* Jacobian F[x][r] = d x / d xi_r = (X_{r+1} - X_0)/2 */
    double Fm[9], Fi[9];
    for (int x = 0; x < dim; ++x)
      for (int r = 0; r < dim; ++r) Fm[x * dim + r] = 0.5 * (X[r + 1][x] - X[0][x]);
    double J;
    if (dim == 2) J = Fm[0] * Fm[3] - Fm[1] * Fm[2];
    else J = Fm[0] * (Fm[4] * Fm[8] - Fm[5] * Fm[7]) - Fm[1] * (Fm[3] * Fm[8] - Fm[5] * Fm[6]) + Fm[2] * (Fm[3] * Fm[7] - Fm[4] * Fm[6]);
    inv_dense(Fm, dim, Fi); /* Fi[r][x] = d xi_r / d x */
    double* g = o->geof + e * ng;
    for (int i = 0; i < dim * dim; ++i) g[i] = Fi[i];
    g[dim * dim] = J;
    for (int f = 0; f < Nf; ++f) {
      const int* fv = dim == 2 ? FV2[f] : FV3[f];
      int opp = 0;
      for (int v = 0; v < nv; ++v) { int on = 0; for (int k = 0; k < dim; ++k) on |= fv[k] == v; if (!on) opp = v; }
      double n[3], meas;
      if (dim == 2) {
        double t[2]; vsub(X[fv[1]], X[fv[0]], t, 2);
        meas = sqrt(t[0] * t[0] + t[1] * t[1]);
        n[0] = t[1] / meas; n[1] = -t[0] / meas;
      } else {
        double a[3], b[3]; vsub(X[fv[1]], X[fv[0]], a, 3); vsub(X[fv[2]], X[fv[0]], b, 3);
        n[0] = a[1] * b[2] - a[2] * b[1]; n[1] = a[2] * b[0] - a[0] * b[2]; n[2] = a[0] * b[1] - a[1] * b[0];
        const double l = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        meas = 0.5 * l;
        for (int k = 0; k < 3; ++k) n[k] /= l;
      }
      double w[3], dp = 0;
      vsub(X[opp], X[fv[0]], w, dim);
      for (int k = 0; k < dim; ++k) dp += n[k] * w[k];
      if (dp > 0) for (int k = 0; k < dim; ++k) n[k] = -n[k];
      for (int k = 0; k < dim; ++k) o->nrm[(e * Nf + f) * dim + k] = n[k];
      o->sJ[e * Nf + f] = meas / 2.0; /* reference face measure is 2 */
      o->fsc[e * Nf + f] = (meas / 2.0) / J;
    }
  }
  /* faces */
  if (cF >= 0) {
    if (2 * cF + cB != E * Nf) { oracle_destroy(o); return NULL; }
    o->F = cF; o->B = cB;
    o->f2c = malloc(sizeof(int) * 2 * (cF + 1)); o->fnum = malloc(sizeof(int) * 2 * (cF + 1));
    o->b2c = malloc(sizeof(int) * (cB + 1)); o->bnum = malloc(sizeof(int) * (cB + 1)); o->bct = calloc(cB + 1, sizeof(int));
    memcpy(o->f2c, cf2c, sizeof(int) * 2 * cF); memcpy(o->fnum, cfnum, sizeof(int) * 2 * cF);
    memcpy(o->b2c, cb2c, sizeof(int) * cB); memcpy(o->bnum, cbnum, sizeof(int) * cB);
    for (int64_t i = 0; i < 2 * cF; ++i)
      if (o->f2c[i] < 0 || o->f2c[i] >= E || o->fnum[i] < 0 || o->fnum[i] >= Nf) { oracle_destroy(o); return NULL; }
    o->cfb = malloc(sizeof(int) * E * Nf);
    for (int64_t i = 0; i < E * Nf; ++i) o->cfb[i] = -1;
    for (int64_t i = 0; i < cB; ++i) {
      if (o->b2c[i] < 0 || o->b2c[i] >= E || o->bnum[i] < 0 || o->bnum[i] >= Nf) { oracle_destroy(o); return NULL; }
      o->cfb[(int64_t)o->b2c[i] * Nf + o->bnum[i]] = (int)i;
    }
    o->mapLpos = face_map(o, &o->R);
    if (!o->mapLpos) { oracle_destroy(o); return NULL; }
    return o;
  }
  facerec* recs = malloc(sizeof(facerec) * E * Nf);
  for (int64_t e = 0; e < E; ++e)
    for (int f = 0; f < Nf; ++f) {
      facerec* rr = recs + e * Nf + f;
      const int* fv = dim == 2 ? FV2[f] : FV3[f];
      rr->key[0] = rr->key[1] = rr->key[2] = -1;
      for (int k = 0; k < dim; ++k) rr->key[k] = o->c2n[e * nv + fv[k]];
      for (int a = 0; a < dim; ++a)
        for (int b = a + 1; b < dim; ++b)
          if (rr->key[b] < rr->key[a]) { int t = rr->key[a]; rr->key[a] = rr->key[b]; rr->key[b] = t; }
      rr->cell = e; rr->face = f;
    }
  qsort(recs, E * Nf, sizeof(facerec), cmp_face);
  int64_t* inter = malloc(sizeof(int64_t) * 4 * E * Nf);
  int64_t* bnd = malloc(sizeof(int64_t) * 2 * E * Nf);
  int64_t ni = 0, nb = 0;
  for (int64_t i = 0; i < E * Nf;) {
    int64_t j = i + 1;
    while (j < E * Nf && !memcmp(recs[j].key, recs[i].key, sizeof(recs[i].key))) ++j;
    if (j - i == 1) { bnd[2 * nb] = recs[i].cell; bnd[2 * nb + 1] = recs[i].face; ++nb; }
    else { inter[4 * ni] = recs[i].cell; inter[4 * ni + 1] = recs[i].face; inter[4 * ni + 2] = recs[i + 1].cell; inter[4 * ni + 3] = recs[i + 1].face; ++ni; }
    i = j;
  }
  free(recs);
  qsort(inter, ni, sizeof(int64_t) * 4, cmp_pair);
  qsort(bnd, nb, sizeof(int64_t) * 2, cmp_pair);
  o->F = ni; o->B = nb;
  o->f2c = malloc(sizeof(int) * 2 * (ni + 1)); o->fnum = malloc(sizeof(int) * 2 * (ni + 1));
  o->b2c = malloc(sizeof(int) * (nb + 1)); o->bnum = malloc(sizeof(int) * (nb + 1)); o->bct = calloc(nb + 1, sizeof(int));
  for (int64_t i = 0; i < ni; ++i) {
    o->f2c[2 * i] = (int)inter[4 * i]; o->f2c[2 * i + 1] = (int)inter[4 * i + 2];
    o->fnum[2 * i] = (int)inter[4 * i + 1]; o->fnum[2 * i + 1] = (int)inter[4 * i + 3];
  }
  free(o->mapLpos);
  o->mapLpos = face_map(o, &o->R);
  if (!o->mapLpos) { free(inter); free(bnd); oracle_destroy(o); return NULL; }
  o->cfb = malloc(sizeof(int) * E * Nf);
  for (int64_t i = 0; i < E * Nf; ++i) o->cfb[i] = -1;
  for (int64_t i = 0; i < nb; ++i) {
    o->b2c[i] = (int)bnd[2 * i]; o->bnum[i] = (int)bnd[2 * i + 1];
    o->cfb[bnd[2 * i] * Nf + bnd[2 * i + 1]] = (int)i;
  }
  free(inter); free(bnd);
  return o;
}

void oracle_destroy(oracle_t* o) {
  if (!o) return;
  free(o->coords); free(o->c2n); free(o->geof); free(o->nrm); free(o->sJ); free(o->fsc);
  free(o->f2c); free(o->fnum); free(o->b2c); free(o->bnum); free(o->bct);
  free(o->mapLpos); free(o->cfb); free(o->k); free(o->mm);
  free(o);
}

void oracle_set_penalty_order(oracle_t* o, int n) { o->pen_order = n; }
int oracle_np(const oracle_t* o) { return o->R.Np; }
int oracle_nfp(const oracle_t* o) { return o->R.Nfp; }
int64_t oracle_n_faces(const oracle_t* o) { return o->F; }
int64_t oracle_n_bfaces(const oracle_t* o) { return o->B; }
void oracle_bfaces(const oracle_t* o, int* b2c, int* bnum) {
  for (int64_t i = 0; i < o->B; ++i) { b2c[i] = o->b2c[i]; bnum[i] = o->bnum[i]; }
}
void oracle_faces(const oracle_t* o, int* f2c, int* fnum) {
  for (int64_t i = 0; i < 2 * o->F; ++i) { f2c[i] = o->f2c[i]; fnum[i] = o->fnum[i]; }
}
void oracle_set_bc_types(oracle_t* o, const int* t) { for (int64_t i = 0; i < o->B; ++i) o->bct[i] = t[i]; }
void oracle_set_factor(oracle_t* o, const double* k) {
  free(o->k);
  o->k = NULL;
  if (k) {
    o->k = malloc(sizeof(double) * o->E * o->R.Np);
    memcpy(o->k, k, sizeof(double) * o->E * o->R.Np);
  }
}

void oracle_set_mm_factor(oracle_t* o, const double* mm) {
  free(o->mm);
  o->mm = NULL;
  if (mm) {
    o->mm = malloc(sizeof(double) * o->E * o->R.Np);
    memcpy(o->mm, mm, sizeof(double) * o->E * o->R.Np);
  }
}
void oracle_set_mm_const(oracle_t* o, double mm) { o->mmc = mm; }

void oracle_node_coords(const oracle_t* o, double* xyz) {
  const int dim = o->dim, nv = dim + 1, Np = o->R.Np;
  for (int64_t e = 0; e < o->E; ++e)
    for (int i = 0; i < Np; ++i) {
      double lam[4];
      barycentric(dim, o->R.r[i], o->R.s[i], o->R.t[i], lam);
      for (int x = 0; x < dim; ++x) {
        double v = 0;
        for (int k = 0; k < nv; ++k) v += lam[k] * o->coords[(size_t)o->c2n[e * nv + k] * dim + x];
        xyz[((size_t)x * o->E + e) * Np + i] = v;
      }
    }
}

/* ------------------------------------------------------------------ */
/* operator (one order)                                                */
/* ------------------------------------------------------------------ */

/* mat_free_tau_c: fvmf_2d_calc_tau_faces.h:3-13, fvmf_2d_calc_tau_bfaces.h:2-10 */
static void calc_tau(const oracle_t* o, const refel* R, const int* map, const double* k, int N,
                     double* tau) {
  const int Nf = R->Nf, Nfp = R->Nfp, Np = R->Np;
  const double pen = 2.0 * (N + 1) * (N + 2);
  for (int64_t i = 0; i < o->F; ++i) {
    const int L = o->f2c[2 * i], Rc = o->f2c[2 * i + 1], fL = o->fnum[2 * i], fR = o->fnum[2 * i + 1];
    double g = 0;
    for (int j = 0; j < Nfp; ++j) {
      const int iL = R->fmask[fL * Nfp + j];
      const int iR = R->fmask[fR * Nfp + map[i * Nfp + j]];
      const double kL = k ? k[(size_t)L * Np + iL] : 1.0, kR = k ? k[(size_t)Rc * Np + iR] : 1.0;
      g = fmax(g, pen * fmax(o->fsc[L * Nf + fL] * kL, o->fsc[Rc * Nf + fR] * kR));
    }
    tau[L * Nf + fL] = g;
    tau[Rc * Nf + fR] = g;
  }
  for (int64_t i = 0; i < o->B; ++i) {
    const int c = o->b2c[i], f = o->bnum[i];
    double g = k ? k[(size_t)c * Np + R->fmask[f * Nfp]] : 1.0;
    for (int j = 1; j < Nfp; ++j) g = fmax(g, k ? k[(size_t)c * Np + R->fmask[f * Nfp + j]] : 1.0);
    tau[c * Nf + f] = g * pen * o->fsc[c * Nf + f];
  }
}

/* FactorViscousMatrix2D::mult with mass factor 0, one component
 * (src/solvers/2d/slip_matrix/factor_viscous_matrix.cpp:87-209; 3D twin), in three phases:
 * gradients and the interior face buffers, the boundary faces, then flux and assembly. */
typedef struct {
  double *q, *jump, *avg;  /* q[x][e][i], jump[e][f*Nfp+j], avg[x][e][f*Nfp+j] */
} op_bufs;

static op_bufs op_bufs_new(const oracle_t* o, const refel* R) {
  const int NF = R->Nf * R->Nfp;
  op_bufs b;
  b.q = malloc(sizeof(double) * o->E * R->Np * o->dim);
  b.jump = calloc((size_t)o->E * NF, sizeof(double));
  b.avg = calloc((size_t)o->E * NF * o->dim, sizeof(double));
  return b;
}

static void op_bufs_free(op_bufs* b) { free(b->q); free(b->jump); free(b->avg); }

static void op_grad_faces(const oracle_t* o, const refel* R, const int* map, const double* k,
                          const double* in, op_bufs* B) {
  const int dim = o->dim, Np = R->Np, Nfp = R->Nfp, Nf = R->Nf, NF = Nf * Nfp;
  const int64_t E = o->E;
  const int ng = dim * dim + 1;
  double* q = B->q;
  double* jump = B->jump;
  double* avg = B->avg;
  /* op2_gemv DR/DS(/DT) + fvmf_2d_grad: q_x = k (rx ur + sx us + tx ut) */
  OMP_FOR
  for (int64_t e = 0; e < E; ++e) {
    const double* u = in + e * Np;
    const double* g = o->geof + e * ng;
    double ur[3][MAXNP];
    for (int r = 0; r < dim; ++r)
      for (int i = 0; i < Np; ++i) {
        double a = 0;
        for (int j = 0; j < Np; ++j) a += R->D[r][i * Np + j] * u[j];
        ur[r][i] = a;
      }
    for (int i = 0; i < Np; ++i) {
      const double kk = k ? k[e * Np + i] : 1.0;
      for (int x = 0; x < dim; ++x) {
        double a = 0;
        for (int r = 0; r < dim; ++r) a += g[r * dim + x] * ur[r][i];
        q[((size_t)x * E + e) * Np + i] = kk * a;
      }
    }
  }
  /* mesh->jump / mesh->avg over interior faces: [u] = u- - u+, {q} = (q- + q+)/2 */
  OMP_FOR
  for (int64_t f = 0; f < o->F; ++f) {
    const int L = o->f2c[2 * f], Rc = o->f2c[2 * f + 1], fL = o->fnum[2 * f], fR = o->fnum[2 * f + 1];
    for (int j = 0; j < Nfp; ++j) {
      const int iL = R->fmask[fL * Nfp + j];
      const int pR = map[f * Nfp + j];
      const int iR = R->fmask[fR * Nfp + pR];
      const double uL = in[(size_t)L * Np + iL], uR = in[(size_t)Rc * Np + iR];
      jump[(size_t)L * NF + fL * Nfp + j] = uL - uR;
      jump[(size_t)Rc * NF + fR * Nfp + pR] = uR - uL;
      for (int x = 0; x < dim; ++x) {
        const double a = 0.5 * (q[((size_t)x * E + L) * Np + iL] + q[((size_t)x * E + Rc) * Np + iR]);
        avg[((size_t)x * E + L) * NF + fL * Nfp + j] = a;
        avg[((size_t)x * E + Rc) * NF + fR * Nfp + pR] = a;
      }
    }
  }
}

/* vmf_2d_mult_avg_jump.h:29-39 (Dirichlet ghost), Neumann: nothing */
static void op_bc_scalar(const oracle_t* o, const refel* R, const double* in, op_bufs* B) {
  const int dim = o->dim, Np = R->Np, Nfp = R->Nfp, NF = R->Nf * Nfp;
  const int64_t E = o->E;
  const double* q = B->q;
  double* jump = B->jump;
  double* avg = B->avg;
  for (int64_t b = 0; b < o->B; ++b) {
    if (o->bct[b] != 0) continue;
    const int c = o->b2c[b], f = o->bnum[b];
    for (int j = 0; j < Nfp; ++j) {
      const int i = R->fmask[f * Nfp + j];
      jump[(size_t)c * NF + f * Nfp + j] += 2.0 * in[(size_t)c * Np + i];
      for (int x = 0; x < dim; ++x) avg[((size_t)x * E + c) * NF + f * Nfp + j] += q[((size_t)x * E + c) * Np + i];
    }
  }
}

static void op_flux_assemble(const oracle_t* o, const refel* R, const double* k, const double* tau,
                             op_bufs* B, double* out) {
  const int dim = o->dim, Np = R->Np, Nfp = R->Nfp, Nf = R->Nf, NF = Nf * Nfp;
  const int64_t E = o->E;
  const int ng = dim * dim + 1;
  const double* q = B->q;
  double* jump = B->jump;
  double* avg = B->avg;
  /* fvmf_2d_mult_flux.h:5-16 */
  OMP_FOR
  for (int64_t e = 0; e < E; ++e) {
    for (int f = 0; f < Nf; ++f) {
      const double* n = o->nrm + (e * Nf + f) * dim;
      const double sJ = o->sJ[e * Nf + f], t = tau[e * Nf + f];
      for (int j = 0; j < Nfp; ++j) {
        const size_t idx = (size_t)e * NF + f * Nfp + j;
        const double hj = 0.5 * jump[idx];
        double sum = 0;
        for (int x = 0; x < dim; ++x) sum += n[x] * avg[(size_t)x * E * NF + idx];
        jump[idx] = sJ * (t * hj - sum);
        const double fact = k ? k[e * Np + R->fmask[f * Nfp + j]] : 1.0;
        for (int x = 0; x < dim; ++x) avg[(size_t)x * E * NF + idx] = n[x] * sJ * fact * -hj;
      }
    }
  }
  /* mesh->mass(q_x) ; q_x += EMAT avg_x ; out = EMAT jump ; vmf_2d_mult_cells_geof ; out += Dr^T q_r ... */
  OMP_FOR
  for (int64_t e = 0; e < E; ++e) {
    const double* g = o->geof + e * ng;
    const double J = g[dim * dim];
    double qx[3][MAXNP], qr[3][MAXNP];
    for (int x = 0; x < dim; ++x)
      for (int i = 0; i < Np; ++i) {
        double a = 0;
        for (int j = 0; j < Np; ++j) a += R->M[i * Np + j] * q[((size_t)x * E + e) * Np + j];
        a *= J;
        for (int m = 0; m < NF; ++m) a += R->EMAT[i * NF + m] * avg[((size_t)x * E + e) * NF + m];
        qx[x][i] = a;
      }
    double* oe = out + e * Np;
    for (int i = 0; i < Np; ++i) {
      double a = 0;
      for (int m = 0; m < NF; ++m) a += R->EMAT[i * NF + m] * jump[(size_t)e * NF + m];
      oe[i] = a;
    }
    for (int r = 0; r < dim; ++r)
      for (int i = 0; i < Np; ++i) {
        double a = 0;
        for (int x = 0; x < dim; ++x) a += g[r * dim + x] * qx[x][i];
        qr[r][i] = a;
      }
    for (int i = 0; i < Np; ++i) {
      double a = 0;
      for (int r = 0; r < dim; ++r)
        for (int j = 0; j < Np; ++j) a += R->D[r][j * Np + i] * qr[r][j];
      oe[i] += a;
    }
  }
}

static void apply_op(const oracle_t* o, const refel* R, const int* map, const double* k,
                     const double* tau, const double* in, double* out) {
  op_bufs B = op_bufs_new(o, R);
  op_grad_faces(o, R, map, k, in, &B);
  op_bc_scalar(o, R, in, &B);
  op_flux_assemble(o, R, k, tau, &B, out);
  op_bufs_free(&B);
}

/* Mass term of the viscous (Helmholtz) operator on the finest order:
 * per-node factor  out_m += J sum_n mm_n M_mn u_n   (fvmf_2d_mult_mm_geof.h:1-26),
 * constant factor  out_m += mm J sum_n M_mn u_n     (vmf_2d_mult_mm_geof.h:1-26). */
static void mass_term(const oracle_t* o, const double* in, double* out) {
  if (!o->mm && o->mmc == 0.0) return;
  const int Np = o->R.Np, ng = o->dim * o->dim + 1;
  OMP_FOR
  for (int64_t e = 0; e < o->E; ++e) {
    const double J = o->geof[e * ng + o->dim * o->dim];
    const double* u = in + e * Np;
    for (int m = 0; m < Np; ++m) {
      double a = 0;
      for (int n = 0; n < Np; ++n)
        a += (o->mm ? o->mm[e * Np + n] : o->mmc) * o->R.M[m * Np + n] * u[n];
      out[e * Np + m] += a * J;
    }
  }
}

/* Its diagonal: diag_i += mm_i J M_ii (fvmf_2d_mm_diag.h:1-9). */
static void mass_diag(const oracle_t* o, double* diag) {
  if (!o->mm && o->mmc == 0.0) return;
  const int Np = o->R.Np, ng = o->dim * o->dim + 1;
  for (int64_t e = 0; e < o->E; ++e) {
    const double J = o->geof[e * ng + o->dim * o->dim];
    for (int i = 0; i < Np; ++i)
      diag[e * Np + i] += (o->mm ? o->mm[e * Np + i] : o->mmc) * J * o->R.M[i * Np + i];
  }
}

/* normal-derivative matrix rows scaled by k: Dn[i][j] = k_i * n.(grad l_j)(x_i) */
static void normal_dmat(const oracle_t* o, const refel* R, int64_t e, int f, const double* k, double* Dn) {
  const int dim = o->dim, Np = R->Np, Nf = R->Nf, ng = dim * dim + 1;
  const double* g = o->geof + e * ng;
  const double* n = o->nrm + (e * Nf + f) * dim;
  double c[3];
  for (int r = 0; r < dim; ++r) { c[r] = 0; for (int x = 0; x < dim; ++x) c[r] += n[x] * g[r * dim + x]; }
  for (int i = 0; i < Np; ++i)
    for (int j = 0; j < Np; ++j) {
      double a = 0;
      for (int r = 0; r < dim; ++r) a += c[r] * R->D[r][i * Np + j];
      Dn[i * Np + j] = a * (k ? k[e * Np + i] : 1.0);
    }
}

/* fvmf_2d_op1_diag.h + fvmf_2d_op2_diag.h + fvmf_2d_bop_diag.h (3D twins). */
/* bct: bc type per bface (NULL: the oracle's own); nrm_comp >= 0: the diagonal of that
 * component of the vector operator, whose faces of type 2 are slip faces weighted by
 * n_comp^2 (fvmf_2d_bop_diag.h:24-60). */
static void calc_diag_bc(const oracle_t* o, const refel* R, const double* k, const double* tau,
                         int consistent, const int* bct, int nrm_comp, double* diag);
static void calc_diag(const oracle_t* o, const refel* R, const double* k, const double* tau,
                      int consistent, double* diag) {
  calc_diag_bc(o, R, k, tau, consistent, NULL, -1, diag);
}
static void calc_diag_bc(const oracle_t* o, const refel* R, const double* k, const double* tau,
                         int consistent, const int* bct, int nrm_comp, double* diag) {
  if (!bct) bct = o->bct;
  const int dim = o->dim, Np = R->Np, Nf = R->Nf, ng = dim * dim + 1;
  const int64_t E = o->E;
  OMP_FOR
  for (int64_t e = 0; e < E; ++e) {
    const double* g = o->geof + e * ng;
    double* d = diag + e * Np;
    for (int i = 0; i < Np; ++i) d[i] = 0;
    for (int x = 0; x < dim; ++x) {
      double Dx[MAXNP * MAXNP], Dk[MAXNP * MAXNP];
      for (int i = 0; i < Np; ++i)
        for (int j = 0; j < Np; ++j) {
          double a = 0;
          for (int r = 0; r < dim; ++r) a += g[r * dim + x] * R->D[r][i * Np + j];
          Dx[i * Np + j] = a;
          Dk[i * Np + j] = a * (k ? k[e * Np + i] : 1.0);
        }
      for (int i = 0; i < Np; ++i) {
        double s = 0;
        for (int kk = 0; kk < Np; ++kk) {
          double Dt = 0; /* (M Dk)[kk][i] */
          for (int m = 0; m < Np; ++m) Dt += R->M[kk * Np + m] * Dk[m * Np + i];
          s += Dx[kk * Np + i] * Dt;
        }
        d[i] += s;
      }
    }
    for (int i = 0; i < Np; ++i) d[i] *= g[dim * dim];
    /* face self terms (interior: 1/2 sJ (...); Dirichlet: sJ (...); Neumann: none) */
    for (int f = 0; f < Nf; ++f) {
      double scale;
      /* find the face kind for this cell face */
      const int bf = o->cfb[e * Nf + f];
      const int is_b = bf >= 0;
      double w = 1.0;
      if (is_b && bct[bf] != 0) {
        if (!(nrm_comp >= 0 && bct[bf] == 2)) continue;
        const double nc = o->nrm[(e * Nf + f) * dim + nrm_comp];
        w = nc * nc;
      }
      scale = (is_b ? 1.0 : 0.5) * w;
      double Dn[MAXNP * MAXNP];
      normal_dmat(o, R, e, f, k, Dn);
      const double* MF = R->MF[f];
      const double sJ = o->sJ[e * Nf + f], t = tau[e * Nf + f];
      for (int i = 0; i < Np; ++i) {
        double s = 0;
        for (int kk = 0; kk < Np; ++kk) {
          const double second = consistent
              ? (k ? Dn[kk * Np + i] / k[e * Np + kk] * k[e * Np + i] : Dn[kk * Np + i]) * MF[kk * Np + i]
              : Dn[kk * Np + i] * MF[kk * Np + i];
          s += MF[i * Np + kk] * Dn[kk * Np + i] + second;
        }
        d[i] += scale * sJ * (t * MF[i * Np + i] - s);
      }
    }
  }
}

/* The penalty uses the program's order DG_ORDER (fvmf_2d_calc_tau_faces.h:9) on every
 * p-multigrid level, i.e. the finest order; this keeps the rediscretised coarse operators
 * equal to the Galerkin products P^T A P. */
static double* ref_tau(const oracle_t* o, const refel* R, const int* map, const double* k) {
  double* tau = malloc(sizeof(double) * o->E * R->Nf);
  calc_tau(o, R, map, k, o->pen_order, tau);
  return tau;
}

void oracle_mult(oracle_t* o, const double* in, double* out) {
  double* tau = ref_tau(o, &o->R, o->mapLpos, o->k);
  apply_op(o, &o->R, o->mapLpos, o->k, tau, in, out);
  mass_term(o, in, out);
  free(tau);
}

void oracle_diag(oracle_t* o, int consistent, double* out) {
  double* tau = ref_tau(o, &o->R, o->mapLpos, o->k);
  calc_diag(o, &o->R, o->k, tau, consistent, out);
  mass_diag(o, out);
  free(tau);
}

static void apply_bc_bct(oracle_t* o, const int* bct, const double* bc, double* rhs);
/* fvmf_2d_apply_bc.h:24-76 (vmf_2d_apply_bc.h when the factor is 1) */
void oracle_apply_bc(oracle_t* o, const double* bc, double* rhs) { apply_bc_bct(o, o->bct, bc, rhs); }
static void apply_bc_bct(oracle_t* o, const int* bct, const double* bc, double* rhs) {
  const refel* R = &o->R;
  const int Np = R->Np, Nfp = R->Nfp, Nf = R->Nf, N = R->N;
  for (int64_t b = 0; b < o->B; ++b) {
    const int c = o->b2c[b], f = o->bnum[b];
    const int* fm = R->fmask + f * Nfp;
    const double sJ = o->sJ[c * Nf + f];
    const double* g = bc + b * Nfp;
    double* rh = rhs + (size_t)c * Np;
    if (bct[b] == 0) {
      double Dn[MAXNP * MAXNP];
      normal_dmat(o, R, c, f, o->k, Dn);
      double gtau = o->k ? o->k[(size_t)c * Np + fm[0]] : 1.0;
      for (int j = 1; j < Nfp; ++j) gtau = fmax(gtau, o->k ? o->k[(size_t)c * Np + fm[j]] : 1.0);
      gtau *= 2.0 * (N + 1) * (N + 2) * o->fsc[c * Nf + f];
      double tmp[MAXNP];
      for (int i = 0; i < Np; ++i) {
        tmp[i] = 0;
        for (int j = 0; j < Nfp; ++j) {
          const double m = R->MF[f][i * Np + fm[j]];
          rh[i] += gtau * sJ * m * g[j];
          tmp[i] += sJ * m * g[j];
        }
      }
      for (int i = 0; i < Np; ++i)
        for (int j = 0; j < Np; ++j) rh[i] += -Dn[j * Np + i] * tmp[j];
    } else if (bct[b] == 1) {
      for (int i = 0; i < Np; ++i)
        for (int j = 0; j < Nfp; ++j) rh[i] += sJ * R->MF[f][i * Np + fm[j]] * g[j];
    }
  }
}

void oracle_mass(oracle_t* o, const double* in, double* out) {
  const int Np = o->R.Np, ng = o->dim * o->dim + 1;
  OMP_FOR
  for (int64_t e = 0; e < o->E; ++e) {
    const double J = o->geof[e * ng + o->dim * o->dim];
    for (int i = 0; i < Np; ++i) {
      double a = 0;
      for (int j = 0; j < Np; ++j) a += o->R.M[i * Np + j] * in[e * Np + j];
      out[e * Np + i] = J * a;
    }
  }
}

double oracle_l2sq(oracle_t* o, const double* e) {
  const int Np = o->R.Np, ng = o->dim * o->dim + 1;
  double s = 0;
  for (int64_t c = 0; c < o->E; ++c) {
    const double J = o->geof[c * ng + o->dim * o->dim];
    for (int i = 0; i < Np; ++i)
      for (int j = 0; j < Np; ++j) s += J * o->R.M[i * Np + j] * e[c * Np + j] * e[c * Np + j];
  }
  return s;
}

/* Gradient with central flux (DGMesh::grad_with_central_flux of the un-vendored toolkit,
 * called at src/solvers/2d/ins_solver.cpp:337 before project_velocity). Parity unpinned
 * against the toolkit's code; this restates the textbook strong form (Hesthaven & Warburton,
 * Nodal DG Methods, 2008, sec. 7.1 with a central flux):
 *   grad_h p = sum_r grad(xi_r) D_r p + M^-1 sum_f MF_f (fscale_f n_f (p* - p-)),
 *   p* = (p- + p+)/2 on interior faces and p* = p- on boundary faces. */
void oracle_grad_central(oracle_t* o, const double* p, double* gx, double* gy, double* gz) {
  const refel* R = &o->R;
  const int dim = o->dim, Np = R->Np, Nfp = R->Nfp, Nf = R->Nf, NF = Nf * Nfp, ng = dim * dim + 1;
  const int64_t E = o->E;
  double* out[3] = {gx, gy, gz};
  double Minv[MAXNP * MAXNP];
  inv_dense(R->M, Np, Minv);
  double* jump = calloc((size_t)E * NF, sizeof(double)); /* p* - p- = (p+ - p-)/2 */
  for (int64_t f = 0; f < o->F; ++f) {
    const int L = o->f2c[2 * f], Rc = o->f2c[2 * f + 1], fL = o->fnum[2 * f], fR = o->fnum[2 * f + 1];
    for (int j = 0; j < Nfp; ++j) {
      const int iL = R->fmask[fL * Nfp + j];
      const int pR = o->mapLpos[f * Nfp + j];
      const int iR = R->fmask[fR * Nfp + pR];
      const double d = 0.5 * (p[(size_t)Rc * Np + iR] - p[(size_t)L * Np + iL]);
      jump[(size_t)L * NF + fL * Nfp + j] = d;
      jump[(size_t)Rc * NF + fR * Nfp + pR] = -d;
    }
  }
  OMP_FOR
  for (int64_t e = 0; e < E; ++e) {
    const double* g = o->geof + e * ng;
    const double* pe = p + e * Np;
    double pr[3][MAXNP];
    for (int r = 0; r < dim; ++r)
      for (int i = 0; i < Np; ++i) {
        double a = 0;
        for (int j = 0; j < Np; ++j) a += R->D[r][i * Np + j] * pe[j];
        pr[r][i] = a;
      }
    for (int x = 0; x < dim; ++x) {
      double flux[MAXNP];
      for (int i = 0; i < Np; ++i) flux[i] = 0;
      for (int f = 0; f < Nf; ++f) {
        const double c = o->fsc[e * Nf + f] * o->nrm[(e * Nf + f) * dim + x];
        double v[MAXNP];
        for (int i = 0; i < Np; ++i) v[i] = 0;
        for (int j = 0; j < Nfp; ++j) v[R->fmask[f * Nfp + j]] = c * jump[(size_t)e * NF + f * Nfp + j];
        for (int i = 0; i < Np; ++i) {
          double a = 0;
          for (int k = 0; k < Np; ++k) a += R->MF[f][i * Np + k] * v[k];
          flux[i] += a;
        }
      }
      for (int i = 0; i < Np; ++i) {
        double a = 0, b = 0;
        for (int r = 0; r < dim; ++r) a += g[r * dim + x] * pr[r][i];
        for (int k = 0; k < Np; ++k) b += Minv[i * Np + k] * flux[k];
        out[x][e * Np + i] = a + b;
      }
    }
  }
  free(jump);
}

/* ------------------------------------------------------------------ */
/* solvers                                                             */
/* ------------------------------------------------------------------ */

static double vdot(int64_t n, const double* a, const double* b) {
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static) if (g_threads != 1)
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

static void project_mean(int64_t n, double* x) {
  double s = 0;
  for (int64_t i = 0; i < n; ++i) s += x[i];
  s /= (double)n;
  for (int64_t i = 0; i < n; ++i) x[i] -= s;
}

typedef void (*precond_fn)(void* ctx, const double* r, double* z);

/* viscous_solver.cpp:158-266 with the true relative residual as stopping test */
static int fcg(oracle_t* o, const refel* R, const int* map, const double* k, const double* tau, const double* rhs,
               double* x, double rtol, double atol, int max_iter, int nullspace, precond_fn P,
               void* ctx, int* iters, double* relres) {
  const int64_t n = o->E * R->Np;
  double* b = malloc(sizeof(double) * n);
  double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n), *zo = malloc(sizeof(double) * n);
  double *p = malloc(sizeof(double) * n), *Ap = malloc(sizeof(double) * n);
  memcpy(b, rhs, sizeof(double) * n);
  if (nullspace) project_mean(n, b);
  const double bn = sqrt(vdot(n, b, b));
  const double tol = fmax(rtol * bn, atol);
  int it = 0, status = 1;
  double rr = 0;
  const int fine = R->Np == o->R.Np;  /* the mass term belongs to the finest order */
  for (int restart = 0; restart < 50; ++restart) {
    apply_op(o, R, map, k, tau, x, Ap);
    if (fine) mass_term(o, x, Ap);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
    rr = vdot(n, r, r);
    if (!isfinite(rr)) { status = -3; break; }
    if (sqrt(rr) <= tol) { status = 0; break; }
    if (it >= max_iter) break;
    P(ctx, r, z);
    if (nullspace) project_mean(n, z);
    double rz = vdot(n, r, z);
    memcpy(p, z, sizeof(double) * n);
    int bd = 0;
    while (it < max_iter) {
      apply_op(o, R, map, k, tau, p, Ap);
      if (fine) mass_term(o, p, Ap);
      const double pap = vdot(n, p, Ap);
      if (!(pap > 0)) { bd = 1; break; }
      const double alpha = rz / pap;
      for (int64_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
      ++it;
      rr = vdot(n, r, r);
      if (!isfinite(rr)) { bd = 1; break; }
      if (sqrt(rr) <= tol) break;
      double* t = zo; zo = z; z = t;
      P(ctx, r, z);
      if (nullspace) project_mean(n, z);
      const double rzn = vdot(n, r, z);
      double bnum = 0;
      for (int64_t i = 0; i < n; ++i) bnum += r[i] * (z[i] - zo[i]);
      const double beta = rz != 0 ? bnum / rz : 0;
      for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
      rz = rzn;
    }
    if (bd) { status = -3; break; }
  }
  if (status != -3 && nullspace) project_mean(n, x);
  if (iters) *iters = it;
  if (relres) *relres = bn > 0 ? sqrt(rr) / bn : 0;
  free(b); free(r); free(z); free(zo); free(p); free(Ap);
  return status;
}

typedef struct { int64_t n; const double* inv; } jac_ctx;
static void jac_precond(void* c, const double* r, double* z) {
  const jac_ctx* j = c;
  for (int64_t i = 0; i < j->n; ++i) z[i] = j->inv[i] * r[i];
}

int oracle_solve_jacobi(oracle_t* o, const double* rhs, double* x, double rtol, double atol,
                        int max_iter, int nullspace, int* iters, double* relres) {
  const int64_t n = o->E * o->R.Np;
  double* tau = ref_tau(o, &o->R, o->mapLpos, o->k);
  double* d = malloc(sizeof(double) * n);
  calc_diag(o, &o->R, o->k, tau, 1, d);
  mass_diag(o, d);
  for (int64_t i = 0; i < n; ++i) d[i] = 1.0 / d[i];
  jac_ctx c = {n, d};
  const int rc = fcg(o, &o->R, o->mapLpos, o->k, tau, rhs, x, rtol, atol, max_iter, nullspace, jac_precond, &c, iters, relres);
  free(tau); free(d);
  return rc;
}

/* Inverse-mass preconditioner of the viscous solve: z = (J M)^-1 r per cell
 * (mesh->inv_mass, viscous_solver.cpp:334-345), times inv_mass_factor
 * (pre_factor_inv_mass, :347-357) or divided by a per-node factor
 * (pre_recp_dat_factor_inv_mass, :359-366; cg_recp_factor_inv_mass.h). */
typedef struct { const oracle_t* o; const double* Minv; double f; const double* recp; } im_ctx;
static void im_precond(void* vc, const double* r, double* z) {
  const im_ctx* c = vc;
  const oracle_t* o = c->o;
  const int Np = o->R.Np, ng = o->dim * o->dim + 1;
  for (int64_t e = 0; e < o->E; ++e) {
    const double J = o->geof[e * ng + o->dim * o->dim];
    for (int i = 0; i < Np; ++i) {
      double a = 0;
      for (int j = 0; j < Np; ++j) a += c->Minv[i * Np + j] * r[e * Np + j];
      a /= J;
      z[e * Np + i] = c->recp ? a / c->recp[e * Np + i] : a * c->f;
    }
  }
}

int oracle_solve_inv_mass(oracle_t* o, const double* rhs, double* x, double rtol, double atol,
                          int max_iter, double inv_mass_factor, const double* recp_factor,
                          int* iters, double* relres) {
  const int Np = o->R.Np;
  double* tau = ref_tau(o, &o->R, o->mapLpos, o->k);
  double* Minv = malloc(sizeof(double) * Np * Np);
  if (inv_dense(o->R.M, Np, Minv) != 0) { free(tau); free(Minv); return -1; }
  im_ctx c = {o, Minv, inv_mass_factor, recp_factor};
  const int rc = fcg(o, &o->R, o->mapLpos, o->k, tau, rhs, x, rtol, atol, max_iter, 0, im_precond, &c,
                     iters, relres);
  free(tau); free(Minv);
  return rc;
}

/* ---- Vector (velocity) operator with slip walls ----
 * Matrix2Vec / Matrix3Vec: ViscousMatrix2D::mult (src/solvers/2d/slip_matrix/viscous_matrix.cpp:67-178)
 * and FactorViscousMatrix2D::mult (factor_viscous_matrix.cpp:87-209), 3D twins in
 * src/solvers/3d/slip_matrix. Vectors are nc x cells x Np; bct is nc x bfaces. */

/* vmf_2d_mult_avg_jump.h:1-51 / vmf_3d_mult_avg_jump.h: a face whose every component has type 2
 * adds the projections n_a (n.U) (twice, to the jump) and n_a (n.dU/dx) to the averages of
 * each component a; otherwise each component of type 0 takes the Dirichlet ghost and the
 * others nothing. */
static void op_bc_vec(const oracle_t* o, const refel* R, int nc, const int* bct, const double* in,
                      op_bufs* B) {
  const int dim = o->dim, Np = R->Np, Nfp = R->Nfp, Nf = R->Nf, NF = Nf * Nfp;
  const int64_t E = o->E, NB = o->B;
  for (int64_t b = 0; b < NB; ++b) {
    const int c = o->b2c[b], f = o->bnum[b];
    int all2 = 1;
    for (int a = 0; a < nc; ++a) all2 = all2 && bct[a * NB + b] == 2;
    if (all2) {
      const double* n = o->nrm + ((size_t)c * Nf + f) * dim;
      for (int j = 0; j < Nfp; ++j) {
        const int i = R->fmask[f * Nfp + j];
        const size_t ind = (size_t)c * NF + f * Nfp + j;
        double nu = 0, nq[3] = {0, 0, 0};
        for (int d = 0; d < nc; ++d) {
          nu += n[d] * in[((size_t)d * E + c) * Np + i];
          for (int x = 0; x < dim; ++x) nq[x] += n[d] * B[d].q[((size_t)x * E + c) * Np + i];
        }
        for (int a = 0; a < nc; ++a) {
          B[a].jump[ind] += 2.0 * n[a] * nu;
          for (int x = 0; x < dim; ++x) B[a].avg[(size_t)x * E * NF + ind] += n[a] * nq[x];
        }
      }
      continue;
    }
    for (int a = 0; a < nc; ++a) {
      if (bct[a * NB + b] != 0) continue;
      for (int j = 0; j < Nfp; ++j) {
        const int i = R->fmask[f * Nfp + j];
        const size_t ind = (size_t)c * NF + f * Nfp + j;
        B[a].jump[ind] += 2.0 * in[((size_t)a * E + c) * Np + i];
        for (int x = 0; x < dim; ++x) B[a].avg[(size_t)x * E * NF + ind] += B[a].q[((size_t)x * E + c) * Np + i];
      }
    }
  }
}

static void apply_vec(const oracle_t* o, const double* tau, int nc, const int* bct, const double* in,
                      double* out) {
  const refel* R = &o->R;
  const size_t cs = (size_t)o->E * R->Np;
  op_bufs B[3];
  for (int a = 0; a < nc; ++a) {
    B[a] = op_bufs_new(o, R);
    op_grad_faces(o, R, o->mapLpos, o->k, in + a * cs, &B[a]);
  }
  op_bc_vec(o, R, nc, bct, in, B);
  for (int a = 0; a < nc; ++a) {
    op_flux_assemble(o, R, o->k, tau, &B[a], out + a * cs);
    mass_term(o, in + a * cs, out + a * cs);
    op_bufs_free(&B[a]);
  }
}

/* Per-component bc types as the scalar kernels see them: 2 stays only where every component
 * is 2 (a slip wall); a lone 2 becomes 3 (no terms, vmf_2d_apply_bc.h:70-74,122-126). */
static int* vec_types(const oracle_t* o, int nc, const int* bct) {
  const int64_t NB = o->B;
  int* t = malloc(sizeof(int) * (nc * NB + 1));
  for (int64_t b = 0; b < NB; ++b) {
    int all2 = 1;
    for (int a = 0; a < nc; ++a) all2 = all2 && bct[a * NB + b] == 2;
    for (int a = 0; a < nc; ++a) {
      const int v = bct[a * NB + b];
      t[a * NB + b] = v == 2 ? (all2 ? 2 : 3) : v;
    }
  }
  return t;
}

void oracle_mult_vec(oracle_t* o, int nc, const int* bct, const double* in, double* out) {
  double* tau = ref_tau(o, &o->R, o->mapLpos, o->k);
  apply_vec(o, tau, nc, bct, in, out);
  free(tau);
}

/* FactorViscousMatrix2D::calc_diag (factor_viscous_matrix.cpp:211-256): the scalar diagonal of
 * each component with its own bc types, slip faces weighted by n_a^2 (fvmf_2d_bop_diag.h:24-60). */
void oracle_diag_vec(oracle_t* o, int nc, const int* bct, int consistent, double* out) {
  double* tau = ref_tau(o, &o->R, o->mapLpos, o->k);
  int* t = vec_types(o, nc, bct);
  const size_t cs = (size_t)o->E * o->R.Np;
  for (int a = 0; a < nc; ++a) {
    calc_diag_bc(o, &o->R, o->k, tau, consistent, t + a * o->B, a, out + a * cs);
    mass_diag(o, out + a * cs);
  }
  free(t);
  free(tau);
}

/* ViscousMatrix2D::apply_bc (viscous_matrix.cpp:26-43 with vmf_2d_apply_bc.h): each component's
 * Dirichlet or Neumann terms; slip faces add nothing. bc is nc x bfaces x Nfp. */
void oracle_apply_bc_vec(oracle_t* o, int nc, const int* bct, const double* bc, double* rhs) {
  int* t = vec_types(o, nc, bct);
  const size_t cs = (size_t)o->E * o->R.Np, bs = (size_t)o->B * o->R.Nfp;
  for (int a = 0; a < nc; ++a) apply_bc_bct(o, t + a * o->B, bc + a * bs, rhs + a * cs);
  free(t);
}

/* ViscousSolver2D::preconditioned_conjugate_gradient (viscous_solver.cpp:158-266; with no
 * preconditioner z = r, conjugate_gradient :85-156) over all components at once, with the
 * flexible beta and the true residual as the stopping test, as the scalar solves here. */
typedef struct { const oracle_t* o; const double* tau; int nc; const int* bct; } vapply_ctx;
typedef struct { int kind; int nc; im_ctx im; const double* inv; int64_t n; } vpre_ctx;
static void vpre(void* vc, const double* r, double* z) {
  const vpre_ctx* c = vc;
  const size_t cs = (size_t)c->im.o->E * c->im.o->R.Np;
  if (c->kind == 0) {
    memcpy(z, r, sizeof(double) * c->n);
  } else if (c->kind == 4) {
    for (int64_t i = 0; i < c->n; ++i) z[i] = c->inv[i] * r[i];
  } else {
    for (int a = 0; a < c->nc; ++a) im_precond((void*)&c->im, r + a * cs, z + a * cs);
  }
}

int oracle_solve_vec(oracle_t* o, int nc, const int* bct, const double* bc, const double* rhs, double* x,
                     double rtol, double atol, int max_iter, int precond, double inv_mass_factor,
                     const double* recp_factor, int* iters, double* relres) {
  const int Np = o->R.Np;
  const int64_t n = (int64_t)nc * o->E * Np;
  double* tau = ref_tau(o, &o->R, o->mapLpos, o->k);
  double* b = malloc(sizeof(double) * n);
  memcpy(b, rhs, sizeof(double) * n);
  if (bc) oracle_apply_bc_vec(o, nc, bct, bc, b);
  double* Minv = malloc(sizeof(double) * Np * Np);
  if (inv_dense(o->R.M, Np, Minv) != 0) { free(tau); free(b); free(Minv); return -1; }
  double* inv = NULL;
  if (precond == 4) {
    inv = malloc(sizeof(double) * n);
    oracle_diag_vec(o, nc, bct, 1, inv);
    for (int64_t i = 0; i < n; ++i) inv[i] = 1.0 / inv[i];
  }
  vpre_ctx pc = {precond, nc, {o, Minv, precond == 2 ? inv_mass_factor : 1.0, precond == 3 ? recp_factor : NULL},
                 inv, n};
  double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n), *zo = malloc(sizeof(double) * n);
  double *p = malloc(sizeof(double) * n), *Ap = malloc(sizeof(double) * n);
  const double bn = sqrt(vdot(n, b, b));
  const double tol = fmax(rtol * bn, atol);
  int it = 0, status = 1;
  double rr = 0;
  for (int restart = 0; restart < 50; ++restart) {
    apply_vec(o, tau, nc, bct, x, Ap);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
    rr = vdot(n, r, r);
    if (!isfinite(rr)) { status = -3; break; }
    if (sqrt(rr) <= tol) { status = 0; break; }
    if (it >= max_iter) break;
    vpre(&pc, r, z);
    double rz = vdot(n, r, z);
    memcpy(p, z, sizeof(double) * n);
    int bd = 0;
    while (it < max_iter) {
      apply_vec(o, tau, nc, bct, p, Ap);
      const double pap = vdot(n, p, Ap);
      if (!(pap > 0)) { bd = 1; break; }
      const double alpha = rz / pap;
      for (int64_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
      ++it;
      rr = vdot(n, r, r);
      if (!isfinite(rr)) { bd = 1; break; }
      if (sqrt(rr) <= tol) break;
      double* t = zo; zo = z; z = t;
      vpre(&pc, r, z);
      const double rzn = vdot(n, r, z);
      double bnum = 0;
      for (int64_t i = 0; i < n; ++i) bnum += r[i] * (z[i] - zo[i]);
      const double beta = rz != 0 ? bnum / rz : 0;
      for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
      rz = rzn;
    }
    if (bd) { status = -3; break; }
  }
  if (iters) *iters = it;
  if (relres) *relres = bn > 0 ? sqrt(rr) / bn : 0;
  free(b); free(r); free(z); free(zo); free(p); free(Ap); free(Minv); free(inv); free(tau);
  return status;
}

/* ---- Over-integrated variable-coefficient operator ----
 * FactorPoissonMatrixFreeDiagOI{2D,3D} (toolkit, not vendored; selected at
 * src/solvers/2d/mp_ins_solver.cpp:98-100, factors set at :580-600): the SIPG form of the
 * nodal operator with every integral taken by cubature, the coefficient given at the points:
 *   volume   sum_q w_q J k_q grad u . grad phi_i
 *   faces    sum_q w_q sJ [ (tau/2 [u] - {k grad u}.n) phi_i - k^-_q [u]/2 n.grad phi_i ]
 * tau from the nodal factor (fvmf_2d_calc_tau_faces.h:9); Dirichlet ghost as
 * vmf_2d_mult_avg_jump.h:32-38. The points are taken from the caller (the library's rules) and
 * matched across faces by their physical coordinates. */

/* nodal basis (values, reference gradients) of R at a reference point */
static void basis_eval(const refel* R, const double* xr, double* phi, double (*dphi)[MAXNP]) {
  const int dim = R->dim, Np = R->Np;
  double m[MAXNP], dm[3][MAXNP];
  const double x[3] = {(1 + xr[0]) / 2, (1 + xr[1]) / 2, dim == 3 ? (1 + xr[2]) / 2 : 0};
  for (int k = 0; k < Np; ++k) {
    const int* e = R->ex[k];
    m[k] = pow(x[0], e[0]) * pow(x[1], e[1]) * (dim == 3 ? pow(x[2], e[2]) : 1.0);
    for (int d = 0; d < dim; ++d) {
      int f[3] = {e[0], e[1], e[2]};
      double v = 0;
      if (f[d] > 0) {
        v = 0.5 * f[d];
        f[d] -= 1;
        v *= pow(x[0], f[0]) * pow(x[1], f[1]) * (dim == 3 ? pow(x[2], f[2]) : 1.0);
      }
      dm[d][k] = v;
    }
  }
  for (int j = 0; j < Np; ++j) {
    double a = 0;
    for (int k = 0; k < Np; ++k) a += m[k] * R->Vi[k * Np + j];
    phi[j] = a;
    for (int d = 0; d < dim; ++d) {
      double b = 0;
      for (int k = 0; k < Np; ++k) b += dm[d][k] * R->Vi[k * Np + j];
      dphi[d][j] = b;
    }
  }
}

static const double RV2[3][2] = {{-1, -1}, {1, -1}, {-1, 1}};
static const double RV3[4][3] = {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};

/* reference and physical coordinates of face point q of face f of cell e */
static void face_point(const oracle_t* o, int64_t e, int f, const double* b, double* xr, double* xp) {
  const int dim = o->dim;
  for (int x = 0; x < dim; ++x) { xr[x] = 0; xp[x] = 0; }
  for (int k = 0; k < dim; ++k) {
    const int v = dim == 2 ? FV2[f][k] : FV3[f][k];
    const int node = o->c2n[e * (dim + 1) + v];
    for (int x = 0; x < dim; ++x) {
      xr[x] += b[k] * (dim == 2 ? RV2[v][x] : RV3[v][x]);
      xp[x] += b[k] * o->coords[(size_t)node * dim + x];
    }
  }
}

void oracle_mult_oi(oracle_t* o, int nv, const double* vrst, const double* vw, int ns, const double* sbary,
                    const double* sw, const double* kv, const double* ks, const double* in, double* out) {
  const refel* R = &o->R;
  const int dim = o->dim, Np = R->Np, Nf = R->Nf, ng = dim * dim + 1;
  const int64_t E = o->E;
  double* tau = ref_tau(o, R, o->mapLpos, o->k);
  /* physical gradient of u and of every basis function at a reference point of cell e */
  #define OI_GRADS(e, xr, phi, gphys, gu)                                        \
    do {                                                                         \
      double dref_[3][MAXNP];                                                    \
      basis_eval(R, xr, phi, dref_);                                             \
      const double* g_ = o->geof + (e) * ng;                                     \
      for (int x = 0; x < dim; ++x) {                                            \
        (gu)[x] = 0;                                                             \
        for (int j = 0; j < Np; ++j) {                                           \
          double a = 0;                                                          \
          for (int r = 0; r < dim; ++r) a += g_[r * dim + x] * dref_[r][j];      \
          (gphys)[x][j] = a;                                                     \
          (gu)[x] += a * in[(e) * Np + j];                                       \
        }                                                                        \
      }                                                                          \
    } while (0)
  /* volume */
  OMP_FOR
  for (int64_t e = 0; e < E; ++e) {
    const double J = o->geof[e * ng + dim * dim];
    double* oe = out + e * Np;
    for (int i = 0; i < Np; ++i) oe[i] = 0;
    for (int q = 0; q < nv; ++q) {
      double phi[MAXNP], gp[3][MAXNP], gu[3];
      OI_GRADS(e, vrst + (size_t)q * dim, phi, gp, gu);
      const double c = vw[q] * J * kv[e * nv + q];
      for (int i = 0; i < Np; ++i) {
        double a = 0;
        for (int x = 0; x < dim; ++x) a += gp[x][i] * gu[x];
        oe[i] += c * a;
      }
    }
  }
  /* faces: each cell face in turn, the neighbour's value at the physically matching point */
  OMP_FOR
  for (int64_t e = 0; e < E; ++e) {
    double* oe = out + e * Np;
    for (int f = 0; f < Nf; ++f) {
      const int bf = o->cfb[e * Nf + f];
      int64_t en = -1;
      int fnn = -1;
      if (bf < 0) {
        /* interior: find the face record */
        for (int64_t i = 0; i < o->F && en < 0; ++i) {
          if (o->f2c[2 * i] == e && o->fnum[2 * i] == f) { en = o->f2c[2 * i + 1]; fnn = o->fnum[2 * i + 1]; }
          else if (o->f2c[2 * i + 1] == e && o->fnum[2 * i + 1] == f) { en = o->f2c[2 * i]; fnn = o->fnum[2 * i]; }
        }
        if (en < 0) continue;
      } else if (o->bct[bf] != 0) {
        continue; /* Neumann: nothing */
      }
      const double* n = o->nrm + (e * Nf + f) * dim;
      const double sJ = o->sJ[e * Nf + f], t = tau[e * Nf + f];
      /* the neighbour's points (periodic faces: matched up to the translation of the face) */
      double shift[3] = {0, 0, 0};
      if (en >= 0) {
        double cm[3] = {0, 0, 0}, cn[3] = {0, 0, 0};
        for (int k = 0; k < dim; ++k) {
          const int vm = dim == 2 ? FV2[f][k] : FV3[f][k], vn = dim == 2 ? FV2[fnn][k] : FV3[fnn][k];
          for (int x = 0; x < dim; ++x) {
            cm[x] += o->coords[(size_t)o->c2n[e * (dim + 1) + vm] * dim + x] / dim;
            cn[x] += o->coords[(size_t)o->c2n[en * (dim + 1) + vn] * dim + x] / dim;
          }
        }
        for (int x = 0; x < dim; ++x) shift[x] = cn[x] - cm[x];
      }
      for (int q = 0; q < ns; ++q) {
        double xr[3], xp[3], phi[MAXNP], gp[3][MAXNP], gu[3];
        face_point(o, e, f, sbary + (size_t)q * dim, xr, xp);
        OI_GRADS(e, xr, phi, gp, gu);
        double um = 0;
        for (int j = 0; j < Np; ++j) um += phi[j] * in[e * Np + j];
        const double km = ks[(e * Nf + f) * ns + q];
        double dnum = 0;
        for (int x = 0; x < dim; ++x) dnum += n[x] * gu[x];
        double jump, avg;  /* [u] and {k grad u}.n */
        if (en >= 0) {
          int qn = -1;
          double best = 1e300;
          for (int p = 0; p < ns; ++p) {
            double yr[3], yp[3], d2 = 0;
            face_point(o, en, fnn, sbary + (size_t)p * dim, yr, yp);
            for (int x = 0; x < dim; ++x) d2 += (yp[x] - shift[x] - xp[x]) * (yp[x] - shift[x] - xp[x]);
            if (d2 < best) { best = d2; qn = p; }
          }
          double yr[3], yp[3], phn[MAXNP], gun[3];
          face_point(o, en, fnn, sbary + (size_t)qn * dim, yr, yp);
          {
            const int64_t ee = en;
            double dref_[3][MAXNP];
            basis_eval(R, yr, phn, dref_);
            const double* g_ = o->geof + ee * ng;
            for (int x = 0; x < dim; ++x) {
              gun[x] = 0;
              for (int j = 0; j < Np; ++j) {
                double a = 0;
                for (int r = 0; r < dim; ++r) a += g_[r * dim + x] * dref_[r][j];
                gun[x] += a * in[ee * Np + j];
              }
            }
          }
          double up = 0;
          for (int j = 0; j < Np; ++j) up += phn[j] * in[en * Np + j];
          const double kp = ks[(en * Nf + fnn) * ns + qn];
          double dnup = 0;
          for (int x = 0; x < dim; ++x) dnup += n[x] * gun[x];
          jump = um - up;
          avg = 0.5 * (km * dnum + kp * dnup);
        } else {
          jump = 2 * um;
          avg = km * dnum;
        }
        const double w = sw[q] * sJ;
        for (int i = 0; i < Np; ++i) {
          double dphi_n = 0;
          for (int x = 0; x < dim; ++x) dphi_n += n[x] * gp[x][i];
          oe[i] += w * ((0.5 * t * jump - avg) * phi[i] - 0.5 * km * jump * dphi_n);
        }
      }
    }
  }
  #undef OI_GRADS
  mass_term(o, in, out);
  free(tau);
}

/* ---- p-multigrid ---- */

typedef struct {
  oracle_t* o;
  int nl;
  level_t* L;
  const int *pre, *post, *cheb;
  double safety, minf;
  int coarse_iters;
  int nullspace;
  double** xv;  /* per level */
  double** bv;
  double** wv;  /* scratch: Ax / residual */
  double** dv;
} pmg_ctx;

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static void lvl_apply(pmg_ctx* c, int l, const double* x, double* y) {
  apply_op(c->o, &c->L[l].R, c->L[l].map, c->L[l].k, c->L[l].tau, x, y);
}

static void chebyshev(pmg_ctx* c, int l, const double* f, double* x, int zero_init) {
  const int64_t n = c->o->E * c->L[l].R.Np;
  const int K = c->cheb ? (c->cheb[l] > 0 ? c->cheb[l] : 1) : 3;
  const double hi = c->safety * c->L[l].lmax, lo = c->minf * hi;
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
  double rho_old = 1.0 / sigma;
  double *d = c->dv[l], *w = c->wv[l];
  const double* inv = c->L[l].inv;
  if (zero_init) {
    for (int64_t i = 0; i < n; ++i) { d[i] = inv[i] * f[i] / theta; x[i] = d[i]; }
  } else {
    lvl_apply(c, l, x, w);
    for (int64_t i = 0; i < n; ++i) { d[i] = inv[i] * (f[i] - w[i]) / theta; x[i] += d[i]; }
  }
  for (int kk = 2; kk <= K; ++kk) {
    lvl_apply(c, l, x, w);
    const double rho = 1.0 / (2 * sigma - rho_old);
    const double c1 = rho * rho_old, c2 = 2 * rho / delta;
    for (int64_t i = 0; i < n; ++i) { d[i] = c1 * d[i] + c2 * inv[i] * (f[i] - w[i]); x[i] += d[i]; }
    rho_old = rho;
  }
}

static void coarse_cg(pmg_ctx* c, int l) {
  const int64_t n = c->o->E * c->L[l].R.Np;
  double *x = c->xv[l], *b = c->bv[l];
  const double* inv = c->L[l].inv;
  double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n), *Ap = malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; ++i) { x[i] = 0; r[i] = b[i]; z[i] = inv[i] * r[i]; p[i] = z[i]; }
  double rz = vdot(n, r, z);
  for (int it = 0; it < c->coarse_iters; ++it) {
    lvl_apply(c, l, p, Ap);
    const double pap = vdot(n, p, Ap);
    const double alpha = pap != 0 ? rz / pap : 0;
    for (int64_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; z[i] = inv[i] * r[i]; }
    const double rzn = vdot(n, r, z);
    const double beta = rz != 0 ? rzn / rz : 0;
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rz = rzn;
  }
  free(r); free(z); free(p); free(Ap);
}

static void elem_matvec(int64_t E, int rows, int cols, const double* A, const double* x, double* y, int add) {
  OMP_FOR
  for (int64_t e = 0; e < E; ++e)
    for (int i = 0; i < rows; ++i) {
      double a = 0;
      for (int j = 0; j < cols; ++j) a += A[i * cols + j] * x[e * cols + j];
      y[e * rows + i] = add ? y[e * rows + i] + a : a;
    }
}

static void vcycle(pmg_ctx* c, int l) {
  if (l == c->nl - 1) { coarse_cg(c, l); return; }
  const int64_t E = c->o->E;
  const int npf = c->L[l].R.Np, npc = c->L[l + 1].R.Np;
  const int64_t n = E * npf;
  const int pre = c->pre ? c->pre[l] : 1, post = c->post ? c->post[l] : 1;
  double *x = c->xv[l], *b = c->bv[l], *w = c->wv[l];
  if (pre <= 0) memset(x, 0, sizeof(double) * n);
  for (int s = 0; s < pre; ++s) chebyshev(c, l, b, x, s == 0);
  lvl_apply(c, l, x, w);
  for (int64_t i = 0; i < n; ++i) w[i] = b[i] - w[i];
  /* restriction = P^T */
  double* Pt = malloc(sizeof(double) * npf * npc);
  for (int i = 0; i < npf; ++i) for (int j = 0; j < npc; ++j) Pt[j * npf + i] = c->L[l].P[i * npc + j];
  elem_matvec(E, npc, npf, Pt, w, c->bv[l + 1], 0);
  free(Pt);
  vcycle(c, l + 1);
  elem_matvec(E, npf, npc, c->L[l].P, c->xv[l + 1], x, 1);
  for (int s = 0; s < post; ++s) chebyshev(c, l, b, x, 0);
}

static void pmg_precond(void* vc, const double* r, double* z) {
  pmg_ctx* c = vc;
  const int64_t n = c->o->E * c->L[0].R.Np;
  memcpy(c->bv[0], r, sizeof(double) * n);
  vcycle(c, 0);
  memcpy(z, c->xv[0], sizeof(double) * n);
}

int oracle_solve_pmg(oracle_t* o, const int* orders, int nl, const int* pre_it, const int* post_it,
                     const int* cheb_orders, double eig_safety, double cheb_min_factor,
                     int max_iter_eig, int coarse_iters, const double* rhs, double* x, double rtol,
                     double atol, int max_iter, int nullspace, int* iters, double* relres,
                     double* lmax_out) {
  const int64_t E = o->E;
  pmg_ctx c;
  memset(&c, 0, sizeof(c));
  c.o = o; c.nl = nl; c.pre = pre_it; c.post = post_it; c.cheb = cheb_orders;
  c.safety = eig_safety; c.minf = cheb_min_factor; c.coarse_iters = coarse_iters; c.nullspace = nullspace;
  c.L = calloc(nl, sizeof(level_t));
  c.xv = calloc(nl, sizeof(double*)); c.bv = calloc(nl, sizeof(double*));
  c.wv = calloc(nl, sizeof(double*)); c.dv = calloc(nl, sizeof(double*));
  for (int l = 0; l < nl; ++l) {
    level_t* L = &c.L[l];
    L->N = orders[l];
    build_refel(o->dim, L->N, &L->R);
    const int np = L->R.Np;
    if (l == 0) {
      if (o->k) { L->k = malloc(sizeof(double) * E * np); memcpy(L->k, o->k, sizeof(double) * E * np); }
    } else if (o->k) {
      /* factor at the coarser nodes by interpolation of the finer order */
      double* I = malloc(sizeof(double) * np * c.L[l - 1].R.Np);
      interp_mono(o->dim, &L->R, &c.L[l - 1].R, I);
      L->k = malloc(sizeof(double) * E * np);
      elem_matvec(E, np, c.L[l - 1].R.Np, I, c.L[l - 1].k, L->k, 0);
      free(I);
    }
    L->map = face_map(o, &L->R);
    L->tau = ref_tau(o, &L->R, L->map, L->k);
    L->diag = malloc(sizeof(double) * E * np);
    calc_diag(o, &L->R, L->k, L->tau, 1, L->diag);
    L->inv = malloc(sizeof(double) * E * np);
    for (int64_t i = 0; i < E * np; ++i) L->inv[i] = 1.0 / L->diag[i];
    c.xv[l] = calloc(E * np, sizeof(double)); c.bv[l] = calloc(E * np, sizeof(double));
    c.wv[l] = calloc(E * np, sizeof(double)); c.dv[l] = calloc(E * np, sizeof(double));
  }
  for (int l = 0; l + 1 < nl; ++l) {
    c.L[l].P = malloc(sizeof(double) * c.L[l].R.Np * c.L[l + 1].R.Np);
    interp_mono(o->dim, &c.L[l].R, &c.L[l + 1].R, c.L[l].P);
  }
  /* power iteration on D^-1 A (set_rand_vec / rayleigh_quotient / vec_normalise); the
   * start vector uses the same counter hash as the device (index row*E + cell) */
  for (int l = 0; l + 1 < nl; ++l) {
    level_t* L = &c.L[l];
    const int np = L->R.Np;
    const int64_t n = E * np;
    double *v = malloc(sizeof(double) * n), *w = malloc(sizeof(double) * n);
    const uint64_t seed = 0x5eedULL + 977ULL * L->N;
    for (int64_t e = 0; e < E; ++e)
      for (int i = 0; i < np; ++i) {
        const uint64_t idx = (uint64_t)i * (uint64_t)E + (uint64_t)e;
        const uint64_t z = mix64(seed + 0x9E3779B97F4A7C15ull * (idx + 1));
        v[e * np + i] = 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
      }
    double lam = 0;
    for (int it = 0; it < (max_iter_eig > 0 ? max_iter_eig : 1); ++it) {
      lvl_apply(&c, l, v, w);
      for (int64_t i = 0; i < n; ++i) w[i] *= L->inv[i];
      lam = vdot(n, v, w) / vdot(n, v, v);
      const double nrm = sqrt(vdot(n, w, w));
      if (!(nrm > 0)) break;
      for (int64_t i = 0; i < n; ++i) v[i] = w[i] / nrm;
    }
    L->lmax = lam;
    if (lmax_out) lmax_out[l] = lam;
    free(v); free(w);
  }
  const int rc = fcg(o, &c.L[0].R, c.L[0].map, c.L[0].k, c.L[0].tau, rhs, x, rtol, atol, max_iter, nullspace,
                     pmg_precond, &c, iters, relres);
  for (int l = 0; l < nl; ++l) {
    free(c.L[l].tau); free(c.L[l].k); free(c.L[l].diag); free(c.L[l].inv); free(c.L[l].P); free(c.L[l].map);
    free(c.xv[l]); free(c.bv[l]); free(c.wv[l]); free(c.dv[l]);
  }
  free(c.L); free(c.xv); free(c.bv); free(c.wv); free(c.dv);
  return rc;
}
