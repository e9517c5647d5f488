/* The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only): every entry
 * point on a small triangle mesh and a small tetrahedron mesh at orders 1..4 (3D 1..3), mixed
 * boundary types, a variable coefficient and a mass term. Prints "ok"; the sanitizers abort on
 * any memory or UB error. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "ins_oracle.h"

static void square(int n, double** xy, int** c2n, int64_t* nn, int64_t* nc) {
  *nn = (int64_t)(n + 1) * (n + 1);
  *nc = (int64_t)2 * n * n;
  *xy = malloc(sizeof(double) * 2 * *nn);
  *c2n = malloc(sizeof(int) * 3 * *nc);
  for (int j = 0; j <= n; ++j)
    for (int i = 0; i <= n; ++i) {
      const int v = j * (n + 1) + i;
      const int in = i > 0 && i < n && j > 0 && j < n;
      (*xy)[2 * v] = (double)i / n + (in ? 0.03 * sin(7.0 * v) / n : 0);
      (*xy)[2 * v + 1] = (double)j / n + (in ? 0.03 * cos(5.0 * v) / n : 0);
    }
  int c = 0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const int a = j * (n + 1) + i, b = a + 1, d = a + n + 1, e = d + 1;
      int* t = *c2n + 3 * c;
      t[0] = a; t[1] = b; t[2] = e; ++c;
      t = *c2n + 3 * c;
      t[0] = a; t[1] = e; t[2] = d; ++c;
    }
}

static void cube(int n, double** xyz, int** c2n, int64_t* nn, int64_t* nc) {
  static const int kuhn[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  const int m = n + 1;
  *nn = (int64_t)m * m * m;
  *nc = (int64_t)6 * n * n * n;
  *xyz = malloc(sizeof(double) * 3 * *nn);
  *c2n = malloc(sizeof(int) * 4 * *nc);
  for (int k = 0; k < m; ++k)
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) {
        const int v = (k * m + j) * m + i;
        (*xyz)[3 * v] = (double)i / n;
        (*xyz)[3 * v + 1] = (double)j / n;
        (*xyz)[3 * v + 2] = (double)k / n;
      }
  int c = 0;
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        for (int t = 0; t < 6; ++t) {
          int p[3] = {i, j, k};
          int* q = *c2n + 4 * c++;
          q[0] = (p[2] * m + p[1]) * m + p[0];
          for (int s = 0; s < 3; ++s) {
            p[kuhn[t][s]] += 1;
            q[s + 1] = (p[2] * m + p[1]) * m + p[0];
          }
        }
}

static void run(int dim, int N, const double* coords, int64_t nn, const int* c2n, int64_t nc) {
  oracle_t* o = oracle_create(dim, N, nn, coords, nc, c2n);
  if (!o) { fprintf(stderr, "oracle_create failed\n"); exit(1); }
  const int Np = oracle_np(o), Nfp = oracle_nfp(o);
  const int64_t B = oracle_n_bfaces(o), n = nc * Np;
  int* bt = malloc(sizeof(int) * (B + 1));
  for (int64_t b = 0; b < B; ++b) bt[b] = (int)(b % 3 == 0);
  oracle_set_bc_types(o, bt);
  double* k = malloc(sizeof(double) * n);
  double* x = malloc(sizeof(double) * n * 3);
  double* y = malloc(sizeof(double) * n * 3);
  double* bc = malloc(sizeof(double) * (B + 1) * Nfp * 3);
  for (int64_t i = 0; i < n; ++i) k[i] = 1.5 + sin(0.37 * i);
  for (int64_t i = 0; i < 3 * n; ++i) x[i] = cos(0.11 * i);
  for (int64_t i = 0; i < (B + 1) * Nfp * 3; ++i) bc[i] = sin(0.23 * i);
  oracle_set_factor(o, k);
  oracle_set_mm_const(o, 2.0);
  oracle_mult(o, x, y);
  oracle_diag(o, 1, y);
  oracle_diag(o, 0, y);
  oracle_apply_bc(o, bc, y);
  oracle_mass(o, x, y);
  double* gx = malloc(sizeof(double) * n * 3);
  oracle_grad_central(o, x, gx, gx + n, gx + 2 * n);
  (void)oracle_l2sq(o, x);
  int it = 0;
  double rr = 0;
  for (int64_t i = 0; i < n; ++i) y[i] = 0;
  oracle_solve_jacobi(o, x, y, 1e-8, 0, 200, 0, &it, &rr);
  for (int64_t i = 0; i < n; ++i) y[i] = 0;
  oracle_solve_inv_mass(o, x, y, 1e-8, 0, 200, 0.5, NULL, &it, &rr);
  int orders[4], pre[4], post[4], cheb[4];
  for (int l = 0; l < N; ++l) { orders[l] = N - l; pre[l] = post[l] = 1; cheb[l] = 3; }
  double lmax[4];
  for (int64_t i = 0; i < n; ++i) y[i] = 0;
  oracle_solve_pmg(o, orders, N, pre, post, cheb, 1.1, 0.1, 10, 10, x, y, 1e-8, 0, 50, 0, &it, &rr, lmax);
  /* vector form with slip walls */
  int* bv = malloc(sizeof(int) * (B + 1) * dim);
  for (int c = 0; c < dim; ++c)
    for (int64_t b = 0; b < B; ++b) bv[c * B + b] = b % 4 == 0 ? 0 : (b % 4 == 1 ? 1 : 2);
  oracle_mult_vec(o, dim, bv, x, y);
  oracle_diag_vec(o, dim, bv, 1, y);
  oracle_apply_bc_vec(o, dim, bv, bc, y);
  for (int64_t i = 0; i < dim * n; ++i) y[i] = 0;
  oracle_solve_vec(o, dim, bv, bc, x, y, 1e-8, 0, 200, 4, 1.0, NULL, &it, &rr);
  /* over-integration with a simple symmetric rule: the vertices' centroid only */
  const double vr[3] = {dim == 2 ? -1.0 / 3 : -0.5, dim == 2 ? -1.0 / 3 : -0.5, -0.5};
  const double vw = dim == 2 ? 2.0 : 4.0 / 3.0;
  const double sb[3] = {1.0 / dim, 1.0 / dim, 1.0 / dim};
  const double sw = 2.0;
  double* kv = malloc(sizeof(double) * nc);
  double* ks = malloc(sizeof(double) * nc * (dim + 1));
  for (int64_t i = 0; i < nc; ++i) kv[i] = 1.0 + 0.1 * (i % 5);
  for (int64_t i = 0; i < nc * (dim + 1); ++i) ks[i] = 1.0 + 0.1 * (i % 7);
  oracle_mult_oi(o, 1, vr, &vw, 1, sb, &sw, kv, ks, x, y);
  free(kv); free(ks); free(bv); free(gx);
  free(bt); free(k); free(x); free(y); free(bc);
  oracle_destroy(o);
}

int main(void) {
  double* xy;
  int* c2n;
  int64_t nn, nc;
  oracle_set_threads(1);
  square(3, &xy, &c2n, &nn, &nc);
  for (int N = 1; N <= 4; ++N) run(2, N, xy, nn, c2n, nc);
  free(xy); free(c2n);
  cube(2, &xy, &c2n, &nn, &nc);
  for (int N = 1; N <= 3; ++N) run(3, N, xy, nn, c2n, nc);
  free(xy); free(c2n);
  printf("ok\n");
  return 0;
}
