// Batched kernels: one block per 256-cell tile, any problem, any order.
// Grids are sized to a fixed capacity; blocks past the device tile count exit,
// so a captured graph stays valid while problems are admitted and retired.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <type_traits>

#include "batch.h"
#include "reduce.h"
#include "sipg_cell.h"

namespace insp {
namespace batch {

namespace {

inline void check(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw std::runtime_error(std::string("batched kernel launch failed: ") + what + ": " +
                             hipGetErrorString(e));
}

constexpr int kB = kTile;

struct Cell {
  int t, slot;
  int64_t e;
  bool live, valid;
};

// Wave-uniform copies: per-tile values (problem slot, its pointers and sizes) are the same in
// every lane; reading them through readfirstlane keeps them in scalar registers, so the
// operator constants behind them are fetched with scalar loads rather than per lane.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uni(int64_t v) {
  const int lo = __builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const int hi = __builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
template <typename X>
__device__ __forceinline__ X* uni(X* p) {
  return reinterpret_cast<X*>(uni((int64_t)reinterpret_cast<uintptr_t>(p)));
}

// A problem frozen by the device-side convergence mask takes no further work.
__device__ __forceinline__ bool frozen(const BatchProb& P) { return P.scal[BS_STATE] != 0.0; }

template <int TB>
__device__ __forceinline__ Cell locate(const BatchView& b) {
  Cell c;
  c.t = (int)xcd_block();
  c.live = c.t < *b.ntiles;
  c.slot = 0;
  c.e = 0;
  c.valid = false;
  if (c.live) {
    const int2 tt = b.tiles[c.t];
    c.slot = uni(tt.x);
    c.e = uni(tt.y) + threadIdx.x;
    c.valid = c.e < uni(b.P[c.slot].E) && !frozen(b.P[c.slot]);
  }
  return c;
}

template <int TB, int K>
__device__ __forceinline__ void put_partials(const BatchView& b, const Cell& c, const double (&v)[K]) {
  if (!c.live) return;  // uniform per block
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double s = block_sum1<TB>(v[k]);
    if (threadIdx.x == 0) b.partial[(int64_t)k * b.tile_cap + c.t] = s;
  }
}

template <typename T>
__device__ __forceinline__ const T* ctab(const BatchTables& tb, int N) {
  if constexpr (std::is_same<T, float>::value) return tb.C32[N];
  else return tb.C64[N];
}

// Tiles of one order: the block's tile is ooff[NN] + its (XCD-remapped) index.
template <int TB, int NN>
__device__ __forceinline__ Cell locate_order(const BatchView& b) {
  Cell c;
  const int local = (int)xcd_block();
  c.live = local < uni(b.ocount[NN]);
  c.t = uni(b.ooff[NN]) + local;
  c.slot = 0;
  c.e = 0;
  c.valid = false;
  if (c.live) {
    const int2 tt = b.tiles[c.t];
    c.slot = uni(tt.x);
    c.e = uni(tt.y) + threadIdx.x;
    c.valid = c.e < uni(b.P[c.slot].E) && !frozen(b.P[c.slot]);
  }
  return c;
}

// Face traces of order-NN problems (order 1 needs none).
template <int TB, int DIM, typename T, int NN>
__global__ void __launch_bounds__(TB) kb_trace(BatchView b, BatchTables tb, int in, int tr) {
  const Cell c = locate_order<TB, NN>(b);
  if (!c.live) return;  // uniform per block
  const BatchProb& P = b.P[c.slot];
  const int prec = std::is_same<T, float>::value;
  const int64_t S = uni(P.S);
  const T* u = uni((const T*)P.v[in]);
  T* trp = uni((T*)P.v[tr]);
  const T* k = uni((const T*)P.kF[prec]);
  const T* C = uni(ctab<T>(tb, NN));
  const double* Kt = uni(P.Kt);
  if (!c.valid) return;
  trace_cell<DIM, NN, T, 2>(c.e, S, u, Kt, k, as_const(C), trp);
}

// Operator of order-NN problems on their own order (level 0); with LEVEL1 the order-1
// operator of every problem (all tiles), one order per kernel so each is compiled lean.
template <int TB, int DIM, typename T, int NN, bool LEVEL1>
__global__ void __launch_bounds__(TB)
kb_apply(BatchView b, BatchTables tb, int in, int tr, int out, int dot) {
  const Cell c = LEVEL1 ? locate<TB>(b) : locate_order<TB, NN>(b);
  double dv[1] = {0.0};
  if (c.live) {  // uniform per block: the problem's data in scalar registers
    const BatchProb& P = b.P[c.slot];
    const int prec = std::is_same<T, float>::value;
    const int64_t S = uni(P.S);
    const T* u = uni((const T*)P.v[in]);
    T* o = uni((T*)P.v[out]);
    const double* tau = uni(LEVEL1 ? P.tauC : P.tauF);
    const T* k = uni((const T*)(LEVEL1 ? P.kC[prec] : P.kF[prec]));
    const T* C = uni(ctab<T>(tb, NN));
    const int* tab = uni(tb.tab[NN]);
    const double* Kt = uni(P.Kt);
    const int* nbr = uni(P.nbr);
    const int* code = uni(P.code);
    const int* bct = uni(P.bctype);
    if (c.valid) {
      if constexpr (NN == 1) {
        dv[0] = apply_p1_cell<DIM, T, 2>(c.e, S, u, Kt, tau, nbr, code, bct, k, as_const(C), as_const(tab), o);
      } else {
        const T* trp = uni((const T*)P.v[tr]);
        dv[0] = apply_cell<DIM, NN, T, 2>(c.e, S, u, trp, Kt, tau, nbr, code, bct, k, as_const(C), as_const(tab),
                                          o);
      }
    }
  }
  if (dot) put_partials<TB, 1>(b, c, dv);
}

__global__ void __launch_bounds__(kB) kb_finalize(BatchView b, int nvals, int slot0) {
  const int j = blockIdx.x;
  if (j >= *b.nprobs) return;
  const int slot = b.probs[j];
  if (frozen(b.P[slot])) return;  // keeps the values it froze with
  const int2 r = b.ranges[slot];
  for (int k = 0; k < nvals; ++k) {
    double v = 0.0;
    for (int q = threadIdx.x; q < r.y - r.x; q += kB) v += b.partial[(int64_t)k * b.tile_cap + r.x + q];
    v = block_sum1<kB>(v);
    if (threadIdx.x == 0) b.P[slot].scal[slot0 + k] = v;
  }
}

template <typename T>
__device__ __forceinline__ int rows_of(const BatchProb& P, int level) { return level ? P.Np1 : P.Np; }

template <int TB, typename T>
__global__ void __launch_bounds__(TB)
kb_dot(BatchView b, int level, int fa, int fb, int fc, int fd, int nvals) {
  const Cell c = locate<TB>(b);
  double v[2] = {0.0, 0.0};
  if (c.valid) {
    const BatchProb& P = b.P[c.slot];
    const T *a = (const T*)P.v[fa], *bb = (const T*)P.v[fb];
    const T *cc = nvals > 1 ? (const T*)P.v[fc] : nullptr, *dd = nvals > 1 ? (const T*)P.v[fd] : nullptr;
    const int rows = rows_of<T>(P, level);
    for (int i = 0; i < rows; ++i) {
      const int64_t q = i * P.S + c.e;
      v[0] += (double)a[q] * (double)bb[q];
      if (nvals > 1) v[1] += (double)cc[q] * (double)dd[q];
    }
  }
  put_partials<TB, 2>(b, c, v);
}

__device__ __forceinline__ double sdiv(double a, double b) { return b != 0.0 ? a / b : 0.0; }

template <int TB>
__global__ void __launch_bounds__(TB) kb_fcg_update_xr(BatchView b) {
  const Cell c = locate<TB>(b);
  double v[2] = {0.0, 0.0};
  if (c.valid) {
    const BatchProb& P = b.P[c.slot];
    const double alpha = sdiv(P.scal[BS_RZ], P.scal[BS_PAP]);
    double *x = (double*)P.v[BV_X], *r = (double*)P.v[BV_R];
    const double *p = (const double*)P.v[BV_P], *Ap = (const double*)P.v[BV_AP], *z = (const double*)P.v[BV_Z];
    for (int i = 0; i < P.Np; ++i) {
      const int64_t q = i * P.S + c.e;
      x[q] = fma(alpha, p[q], x[q]);
      const double ri = fma(-alpha, Ap[q], r[q]);
      r[q] = ri;
      v[0] += ri * ri;
      v[1] += ri * z[q];
    }
  }
  put_partials<TB, 2>(b, c, v);
}

template <int TB>
__global__ void __launch_bounds__(TB) kb_fcg_update_p(BatchView b) {
  const Cell c = locate<TB>(b);
  if (!c.valid) return;
  const BatchProb& P = b.P[c.slot];
  // flexible (Polak-Ribiere) beta = r.(z - z_old) / (r_old.z_old)  (cg_compute_beta_pre_flex.h)
  const double beta = sdiv(P.scal[BS_RZN] - P.scal[BS_RZO], P.scal[BS_RZ]);
  double* p = (double*)P.v[BV_P];
  const double* z = (const double*)P.v[BV_Z];
  for (int i = 0; i < P.Np; ++i) {
    const int64_t q = i * P.S + c.e;
    p[q] = fma(beta, p[q], z[q]);
  }
}

__global__ void kb_scal_copy(BatchView b, int dst, int src) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= *b.nprobs) return;
  double* s = b.P[b.probs[j]].scal;
  if (s[BS_STATE] != 0.0) return;
  s[dst] = s[src];
}

__global__ void kb_set_state(double* scal, int n, const double* rec, int limits) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double* r = rec + 5 * (int64_t)j;
  double* s = scal + (int64_t)r[0] * BS_COUNT;
  s[BS_STATE] = r[1];
  if (limits) {
    s[BS_ITER] = r[2];
    s[BS_TOL] = r[3];
    s[BS_MAXIT] = r[4];
  }
}

// One thread per problem: count the iteration, then freeze the problem when it broke down,
// reached its tolerance (recursive residual) or its iteration limit.
__global__ void kb_mask(BatchView b) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= *b.nprobs) return;
  double* s = b.P[b.probs[j]].scal;
  if (s[BS_STATE] != 0.0) return;
  const double it = s[BS_ITER] + 1.0;
  s[BS_ITER] = it;
  const double pap = s[BS_PAP], rr = s[BS_RR];
  double st = BST_RUNNING;
  if (!(pap > 0.0) || !isfinite(rr) || !isfinite(pap)) st = BST_BREAKDOWN;
  else if (sqrt(rr) <= s[BS_TOL]) st = BST_TOL;
  else if (it >= s[BS_MAXIT]) st = BST_MAXIT;
  s[BS_STATE] = st;
}

template <int TB>
__global__ void __launch_bounds__(TB) kb_residual(BatchView b) {
  const Cell c = locate<TB>(b);
  double v[1] = {0.0};
  if (c.valid) {
    const BatchProb& P = b.P[c.slot];
    const double *bb = (const double*)P.v[BV_B], *Ap = (const double*)P.v[BV_AP];
    double* r = (double*)P.v[BV_R];
    for (int i = 0; i < P.Np; ++i) {
      const int64_t q = i * P.S + c.e;
      const double t = bb[q] - Ap[q];
      r[q] = t;
      v[0] += t * t;
    }
  }
  put_partials<TB, 1>(b, c, v);
}

template <int TB, typename Tin, typename Tout>
__global__ void __launch_bounds__(TB) kb_convert(BatchView b, int src, int dst) {
  const Cell c = locate<TB>(b);
  if (!c.valid) return;
  const BatchProb& P = b.P[c.slot];
  const Tin* x = (const Tin*)P.v[src];
  Tout* y = (Tout*)P.v[dst];
  for (int i = 0; i < P.Np; ++i) y[i * P.S + c.e] = (Tout)x[i * P.S + c.e];
}

template <int TB, typename T>
__global__ void __launch_bounds__(TB) kb_cheb(BatchView b, int level, int step, int zero_init) {
  const Cell c = locate<TB>(b);
  if (!c.valid) return;
  const BatchProb& P = b.P[c.slot];
  const int prec = std::is_same<T, float>::value;
  const int rows = level ? P.Np1 : P.Np;
  const T* inv = (const T*)(level ? P.invC[prec] : P.invF[prec]);
  const T *f = (const T*)P.v[level ? BV_CB : BV_VB], *ax = (const T*)P.v[level ? BV_CAP : BV_VAX];
  T *d = (T*)P.v[level ? BV_CZ : BV_VD], *x = (T*)P.v[level ? BV_CX : BV_VX];
  const T c1 = (T)(level ? P.amg.oc1[step] : P.c1[step]), c2 = (T)(level ? P.amg.oc2[step] : P.c2[step]);
  for (int i = 0; i < rows; ++i) {
    const int64_t q = i * P.S + c.e;
    const T res = inv[q] * (zero_init ? f[q] : f[q] - ax[q]);
    const T di = step == 0 ? c2 * res : fma(c1, d[q], c2 * res);
    d[q] = di;
    x[q] = zero_init ? di : x[q] + di;
  }
}

template <int TB, typename T>
__global__ void __launch_bounds__(TB) kb_sub_from_b(BatchView b, int level) {
  const Cell c = locate<TB>(b);
  if (!c.valid) return;
  const BatchProb& P = b.P[c.slot];
  const int rows = level ? P.Np1 : P.Np;
  const T* f = (const T*)P.v[level ? BV_CB : BV_VB];
  T* ax = (T*)P.v[level ? BV_CAP : BV_VAX];
  for (int i = 0; i < rows; ++i) {
    const int64_t q = i * P.S + c.e;
    ax[q] = f[q] - ax[q];
  }
}

// ---- vertex-space AMG ----------------------------------------------------

struct Row {
  int slot, r;
  bool valid;
};

// Rows of AMG level k: one 256-row tile per block, from the level's tile table.
__device__ __forceinline__ Row locate_row(const BatchView& b, int k) {
  Row c;
  const int t = (int)xcd_block();
  c.slot = 0;
  c.r = 0;
  c.valid = false;
  if (t < b.lcount[k]) {
    const int2 tt = b.ltiles[(int64_t)k * b.ltile_cap + t];
    c.slot = tt.x;
    c.r = tt.y + threadIdx.x;
    c.valid = c.r < b.P[c.slot].amg.lev[k].n && !frozen(b.P[c.slot]);
  }
  return c;
}

template <typename T>
__global__ void __launch_bounds__(kB) kb_vertex_restrict(BatchView b) {
  const Row c = locate_row(b, 0);
  if (!c.valid) return;
  const BatchProb& P = b.P[c.slot];
  const T* r = (const T*)P.v[BV_CAP];
  T s = 0;
  for (int q = P.amg.vptr[c.r]; q < P.amg.vptr[c.r + 1]; ++q) s += r[P.amg.vlist[q]];
  ((T*)P.amg.lev[0].b)[c.r] = s;
}

template <int TB, typename T>
__global__ void __launch_bounds__(TB) kb_vertex_prolong(BatchView b) {
  const Cell c = locate<TB>(b);
  if (!c.valid) return;
  const BatchProb& P = b.P[c.slot];
  if (P.amg.nlev == 0) return;
  const T* xv = (const T*)P.amg.lev[0].x;
  T* x = (T*)P.v[BV_CX];
  for (int i = 0; i < P.Np1; ++i) {
    const int64_t q = i * P.S + c.e;
    x[q] += xv[P.amg.cv[q]];
  }
}

// CSR products over a 256-row tile of AMG level k: groups of kCsrLanes adjacent lanes share a
// row (lane l takes the row's nonzeros l, l + kCsrLanes, ...) and add their partial sums by
// shuffles in a fixed order; each group walks kCsrLanes rows of the tile. Returns the tile's
// problem slot and first row, or slot -1 when the block has no live tile.
constexpr int kCsrLanes = 8;
__device__ __forceinline__ int2 locate_tile(const BatchView& b, int k) {
  const int t = (int)xcd_block();
  if (t >= b.lcount[k]) return make_int2(-1, 0);
  const int2 tt = b.ltiles[(int64_t)k * b.ltile_cap + t];
  return frozen(b.P[tt.x]) ? make_int2(-1, 0) : tt;
}

template <typename T>
__device__ __forceinline__ T csr_row_dot(const int* __restrict__ ptr, const int* __restrict__ col,
                                         const T* __restrict__ val, const T* __restrict__ x, int r,
                                         bool live) {
  const int l = (int)threadIdx.x % kCsrLanes;
  T s = 0;
  if (live)
    for (int q = ptr[r] + l; q < ptr[r + 1]; q += kCsrLanes) s = fma(val[q], x[col[q]], s);
#pragma unroll
  for (int o = kCsrLanes / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kCsrLanes);
  return s;
}

template <typename T>
__global__ void __launch_bounds__(kB) kb_amg_spmv(BatchView b, int k, int residual) {
  const int2 tt = locate_tile(b, k);
  if (tt.x < 0) return;  // the whole block
  const BatchProb& P = b.P[tt.x];
  if (k >= P.amg.nlev - 1) return;  // the last level is solved densely
  const BatchAmgLevel& L = P.amg.lev[k];
  const T* val = (const T*)L.Aval;
  const T* x = (const T*)L.x;
  constexpr int G = kB / kCsrLanes;
  const int g = (int)threadIdx.x / kCsrLanes;
  for (int i = 0; i < kCsrLanes; ++i) {
    const int r = tt.y + g + G * i;
    const bool live = r < L.n;
    const T s = csr_row_dot<T>(L.Aptr, L.Acol, val, x, r, live);
    if (live && threadIdx.x % kCsrLanes == 0) ((T*)L.r)[r] = residual ? ((const T*)L.b)[r] - s : s;
  }
}

template <typename T>
__global__ void __launch_bounds__(kB) kb_amg_cheb(BatchView b, int k, int step, int zero_init) {
  const Row c = locate_row(b, k);
  if (!c.valid) return;
  const BatchProb& P = b.P[c.slot];
  if (k >= P.amg.nlev - 1) return;
  const BatchAmgLevel& L = P.amg.lev[k];
  const T f = ((const T*)L.b)[c.r];
  const T res = ((const T*)L.inv)[c.r] * (zero_init ? f : f - ((const T*)L.r)[c.r]);
  T* d = (T*)L.d;
  T* x = (T*)L.x;
  const T di = step == 0 ? (T)L.ch2[0] * res : fma((T)L.ch1[step], d[c.r], (T)L.ch2[step] * res);
  d[c.r] = di;
  x[c.r] = zero_init ? di : x[c.r] + di;
}

// Rows of level k+1: b_{k+1} = R_k r_k.
template <typename T>
__global__ void __launch_bounds__(kB) kb_amg_restrict(BatchView b, int k) {
  const int2 tt = locate_tile(b, k + 1);
  if (tt.x < 0) return;  // the whole block
  const BatchProb& P = b.P[tt.x];
  const BatchAmgLevel& L = P.amg.lev[k];
  const int nc = P.amg.lev[k + 1].n;
  const T* val = (const T*)L.Rval;
  const T* res = (const T*)L.r;
  constexpr int G = kB / kCsrLanes;
  const int g = (int)threadIdx.x / kCsrLanes;
  for (int i = 0; i < kCsrLanes; ++i) {
    const int r = tt.y + g + G * i;
    const bool live = r < nc;
    const T s = csr_row_dot<T>(L.Rptr, L.Rcol, val, res, r, live);
    if (live && threadIdx.x % kCsrLanes == 0) ((T*)P.amg.lev[k + 1].b)[r] = s;
  }
}

// Rows of level k: x_k += P_k x_{k+1}.
template <typename T>
__global__ void __launch_bounds__(kB) kb_amg_prolong(BatchView b, int k) {
  const Row c = locate_row(b, k);
  if (!c.valid) return;
  const BatchProb& P = b.P[c.slot];
  if (k >= P.amg.nlev - 1) return;
  const BatchAmgLevel& L = P.amg.lev[k];
  const T* val = (const T*)L.Pval;
  const T* xc = (const T*)P.amg.lev[k + 1].x;
  T s = 0;
  for (int q = L.Pptr[c.r]; q < L.Pptr[c.r + 1]; ++q) s = fma(val[q], xc[L.Pcol[q]], s);
  ((T*)L.x)[c.r] += s;
}

// One block per problem whose last level is k: x = Cinv b (Cinv symmetric, n <= 256).
template <typename T>
__global__ void __launch_bounds__(kB) kb_amg_dense(BatchView b, int k) {
  const int j = blockIdx.x;
  if (j >= *b.nprobs) return;
  const BatchProb& P = b.P[b.probs[j]];
  if (P.amg.nlev == 0 || P.amg.nlev - 1 != k || frozen(P)) return;
  const int n = P.amg.nc, i = threadIdx.x;
  if (i >= n) return;
  const T* M = (const T*)P.amg.cinv;
  const T* bb = (const T*)P.amg.lev[k].b;
  T s = 0;
  for (int q = 0; q < n; ++q) s = fma(M[(int64_t)q * n + i], bb[q], s);
  ((T*)P.amg.lev[k].x)[i] = s;
}

template <typename T, int R, int K, bool ADD>
__device__ __forceinline__ void cell_matvec(const T* __restrict__ A, const T* __restrict__ x,
                                            T* __restrict__ y, int64_t S, int64_t e) {
  T xv[K];
#pragma unroll
  for (int j = 0; j < K; ++j) xv[j] = x[j * S + e];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    T a = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) a = fma(A[i * K + j], xv[j], a);
    y[i * S + e] = ADD ? y[i * S + e] + a : a;
  }
}

template <typename T>
__device__ __forceinline__ const T* ptab(const BatchTables& tb, int N, bool restrict_op) {
  if constexpr (std::is_same<T, float>::value) return restrict_op ? tb.Pt32[N] : tb.P32[N];
  else return restrict_op ? tb.Pt64[N] : tb.P64[N];
}

template <int TB, int DIM, typename T>
__global__ void __launch_bounds__(TB) kb_restrict(BatchView b, BatchTables tb) {
  const Cell c = locate<TB>(b);
  if (!c.valid) return;
  const BatchProb& P = b.P[c.slot];
  const T* r = (const T*)P.v[BV_VAX];
  T* cb = (T*)P.v[BV_CB];
  constexpr int n1 = DIM + 1;
  constexpr int n2 = DIM == 2 ? 6 : 10, n3 = DIM == 2 ? 10 : 20, n4 = DIM == 2 ? 15 : 35;
  const T* A = ptab<T>(tb, P.N, true);
  switch (P.N) {
    case 1: for (int i = 0; i < n1; ++i) cb[i * P.S + c.e] = r[i * P.S + c.e]; break;
    case 2: cell_matvec<T, n1, n2, false>(A, r, cb, P.S, c.e); break;
    case 3: cell_matvec<T, n1, n3, false>(A, r, cb, P.S, c.e); break;
    case 4: cell_matvec<T, n1, n4, false>(A, r, cb, P.S, c.e); break;
    default: break;
  }
}

template <int TB, int DIM, typename T>
__global__ void __launch_bounds__(TB) kb_prolong(BatchView b, BatchTables tb) {
  const Cell c = locate<TB>(b);
  if (!c.valid) return;
  const BatchProb& P = b.P[c.slot];
  const T* cx = (const T*)P.v[BV_CX];
  T* x = (T*)P.v[BV_VX];
  constexpr int n1 = DIM + 1;
  constexpr int n2 = DIM == 2 ? 6 : 10, n3 = DIM == 2 ? 10 : 20, n4 = DIM == 2 ? 15 : 35;
  const T* A = ptab<T>(tb, P.N, false);
  switch (P.N) {
    case 1: for (int i = 0; i < n1; ++i) x[i * P.S + c.e] += cx[i * P.S + c.e]; break;
    case 2: cell_matvec<T, n2, n1, true>(A, cx, x, P.S, c.e); break;
    case 3: cell_matvec<T, n3, n1, true>(A, cx, x, P.S, c.e); break;
    case 4: cell_matvec<T, n4, n1, true>(A, cx, x, P.S, c.e); break;
    default: break;
  }
}

template <int TB, typename T>
__global__ void __launch_bounds__(TB) kb_cg_init(BatchView b) {
  const Cell c = locate<TB>(b);
  double v[1] = {0.0};
  if (c.valid) {
    const BatchProb& P = b.P[c.slot];
    const int prec = std::is_same<T, float>::value;
    const T* inv = (const T*)P.invC[prec];
    const T* cb = (const T*)P.v[BV_CB];
    T *x = (T*)P.v[BV_CX], *r = (T*)P.v[BV_CR], *z = (T*)P.v[BV_CZ], *p = (T*)P.v[BV_CP];
    for (int i = 0; i < P.Np1; ++i) {
      const int64_t q = i * P.S + c.e;
      const T ri = cb[q], zi = inv[q] * ri;
      x[q] = 0;
      r[q] = ri;
      z[q] = zi;
      p[q] = zi;
      v[0] += (double)ri * (double)zi;
    }
  }
  put_partials<TB, 1>(b, c, v);
}

template <int TB, typename T>
__global__ void __launch_bounds__(TB) kb_cg_xr(BatchView b, int num) {
  const Cell c = locate<TB>(b);
  double v[1] = {0.0};
  if (c.valid) {
    const BatchProb& P = b.P[c.slot];
    const int prec = std::is_same<T, float>::value;
    const T a = (T)sdiv(P.scal[num], P.scal[BS_CPAP]);
    const T* inv = (const T*)P.invC[prec];
    const T *p = (const T*)P.v[BV_CP], *Ap = (const T*)P.v[BV_CAP];
    T *x = (T*)P.v[BV_CX], *r = (T*)P.v[BV_CR], *z = (T*)P.v[BV_CZ];
    for (int i = 0; i < P.Np1; ++i) {
      const int64_t q = i * P.S + c.e;
      x[q] = fma(a, p[q], x[q]);
      const T ri = fma(-a, Ap[q], r[q]), zi = inv[q] * ri;
      r[q] = ri;
      z[q] = zi;
      v[0] += (double)ri * (double)zi;
    }
  }
  put_partials<TB, 1>(b, c, v);
}

template <int TB, typename T>
__global__ void __launch_bounds__(TB) kb_cg_p(BatchView b, int num, int den) {
  const Cell c = locate<TB>(b);
  if (!c.valid) return;
  const BatchProb& P = b.P[c.slot];
  const T beta = (T)sdiv(P.scal[num], P.scal[den]);
  const T* z = (const T*)P.v[BV_CZ];
  T* p = (T*)P.v[BV_CP];
  for (int i = 0; i < P.Np1; ++i) {
    const int64_t q = i * P.S + c.e;
    p[q] = fma(beta, p[q], z[q]);
  }
}

template <int TB>
__global__ void __launch_bounds__(TB) kb_sum_d(BatchView b, int field) {
  const Cell c = locate<TB>(b);
  double v[1] = {0.0};
  if (c.valid) {
    const BatchProb& P = b.P[c.slot];
    const double* x = (const double*)P.v[field];
    for (int i = 0; i < P.Np; ++i) v[0] += x[i * P.S + c.e];
  }
  put_partials<TB, 1>(b, c, v);
}

template <int TB>
__global__ void __launch_bounds__(TB) kb_sub_mean_d(BatchView b, int field) {
  const Cell c = locate<TB>(b);
  if (!c.valid) return;
  const BatchProb& P = b.P[c.slot];
  if (!P.nullspace) return;
  const double m = P.scal[BS_MEAN] / P.dofs;
  double* x = (double*)P.v[field];
  for (int i = 0; i < P.Np; ++i) x[i * P.S + c.e] -= m;
}

template <int TB>
__global__ void __launch_bounds__(TB) kb_rand_d(BatchView b, int level, int field) {
  const Cell c = locate<TB>(b);
  if (!c.valid) return;
  const BatchProb& P = b.P[c.slot];
  double* x = (double*)P.v[field];
  const int rows = level ? P.Np1 : P.Np;
  const int N = level ? 1 : P.N;
  for (int i = 0; i < rows; ++i) {
    const uint64_t idx = (uint64_t)(i * P.S + c.e);
    uint64_t z = 0x5eedULL + 977ULL * N + 0x9E3779B97F4A7C15ull * (idx + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    x[i * P.S + c.e] = 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
  }
}

template <int TB>
__global__ void __launch_bounds__(TB) kb_inv_apply_d(BatchView b, int level, int field) {
  const Cell c = locate<TB>(b);
  if (!c.valid) return;
  const BatchProb& P = b.P[c.slot];
  const double* inv = (const double*)(level ? P.invC[0] : P.invF[0]);
  double* x = (double*)P.v[field];
  const int rows = level ? P.Np1 : P.Np;
  for (int i = 0; i < rows; ++i) x[i * P.S + c.e] *= inv[i * P.S + c.e];
}

template <int TB>
__global__ void __launch_bounds__(TB) kb_scale_d(BatchView b, int field, int slot) {
  const Cell c = locate<TB>(b);
  if (!c.valid) return;
  const BatchProb& P = b.P[c.slot];
  const double s = P.scal[slot] > 0 ? 1.0 / sqrt(P.scal[slot]) : 0.0;
  double* x = (double*)P.v[field];
  for (int i = 0; i < P.Np; ++i) x[i * P.S + c.e] *= s;
}

dim3 tiles_grid(const BatchView& b) { return dim3((unsigned)b.tile_cap); }

}  // namespace

#define INSP_TB(...)                                          \
  do {                                                        \
    switch (b.tile) {                                         \
      case 64: { constexpr int TB_ = 64; __VA_ARGS__; } break;   \
      case 128: { constexpr int TB_ = 128; __VA_ARGS__; } break; \
      default: { constexpr int TB_ = 256; __VA_ARGS__; } break;  \
    }                                                         \
  } while (0)

#define INSP_DIMT(CALL)                     \
  do {                                      \
    if (b.dim == 2) { constexpr int D_ = 2; CALL; } \
    else { constexpr int D_ = 3; CALL; }    \
  } while (0)

template <typename T>
void trace(const BatchView& b, const BatchTables& t, int in, int tr, hipStream_t s) {
#define INSP_TRACE(NN)                                                                              \
  INSP_DIMT(INSP_TB(hipLaunchKernelGGL((kb_trace<TB_, D_, T, NN>), dim3((unsigned)b.ocap[NN]), dim3(TB_), \
                                       0, s, b, t, in, tr)))
  INSP_TRACE(2);
  INSP_TRACE(3);
  INSP_TRACE(4);
#undef INSP_TRACE
  check("kb_trace");
}

template <typename T>
void apply(const BatchView& b, const BatchTables& t, int level, int in, int tr, int out, bool dot,
           hipStream_t s) {
  if (level) {
    INSP_DIMT(INSP_TB(hipLaunchKernelGGL((kb_apply<TB_, D_, T, 1, true>), tiles_grid(b), dim3(TB_), 0, s, b,
                                         t, in, tr, out, (int)dot)));
  } else {
#define INSP_APPLY(NN)                                                                              \
  INSP_DIMT(INSP_TB(hipLaunchKernelGGL((kb_apply<TB_, D_, T, NN, false>), dim3((unsigned)b.ocap[NN]),      \
                                       dim3(TB_), 0, s, b, t, in, tr, out, (int)dot)))
    INSP_APPLY(1);
    INSP_APPLY(2);
    INSP_APPLY(3);
    INSP_APPLY(4);
#undef INSP_APPLY
  }
  check("kb_apply");
}

void set_state(double* scal, int n, const double* rec, bool limits, hipStream_t s) {
  hipLaunchKernelGGL(kb_set_state, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, scal, n, rec,
                     (int)limits);
  check("kb_set_state");
}

void mask(const BatchView& b, hipStream_t s) {
  hipLaunchKernelGGL(kb_mask, dim3((unsigned)((b.prob_cap + 63) / 64)), dim3(64), 0, s, b);
  check("kb_mask");
}

void finalize(const BatchView& b, int nvals, int slot, hipStream_t s) {
  hipLaunchKernelGGL(kb_finalize, dim3((unsigned)b.prob_cap), dim3(kB), 0, s, b, nvals, slot);
  check("kb_finalize");
}

template <typename T>
void dot(const BatchView& b, int level, int a, int bb, int c, int d, int nvals, hipStream_t s) {
  INSP_TB(hipLaunchKernelGGL((kb_dot<TB_, T>), tiles_grid(b), dim3(TB_), 0, s, b, level, a, bb, c, d, nvals));
  check("kb_dot");
}

void fcg_update_xr(const BatchView& b, hipStream_t s) {
  INSP_TB(hipLaunchKernelGGL((kb_fcg_update_xr<TB_>), tiles_grid(b), dim3(TB_), 0, s, b));
  check("kb_fcg_update_xr");
}

void fcg_update_p(const BatchView& b, hipStream_t s) {
  INSP_TB(hipLaunchKernelGGL((kb_fcg_update_p<TB_>), tiles_grid(b), dim3(TB_), 0, s, b));
  check("kb_fcg_update_p");
}

void scal_copy(const BatchView& b, int dst, int src, hipStream_t s) {
  hipLaunchKernelGGL(kb_scal_copy, dim3((unsigned)((b.prob_cap + 63) / 64)), dim3(64), 0, s, b, dst, src);
  check("kb_scal_copy");
}

void residual(const BatchView& b, hipStream_t s) {
  INSP_TB(hipLaunchKernelGGL((kb_residual<TB_>), tiles_grid(b), dim3(TB_), 0, s, b));
  check("kb_residual");
}

void to_t(const BatchView& b, bool fp32, int src_d, int dst, hipStream_t s) {
  if (fp32) INSP_TB(hipLaunchKernelGGL((kb_convert<TB_, double, float>), tiles_grid(b), dim3(TB_), 0, s, b, src_d, dst));
  else INSP_TB(hipLaunchKernelGGL((kb_convert<TB_, double, double>), tiles_grid(b), dim3(TB_), 0, s, b, src_d, dst));
  check("kb_convert");
}

void from_t(const BatchView& b, bool fp32, int src, int dst_d, hipStream_t s) {
  if (fp32) INSP_TB(hipLaunchKernelGGL((kb_convert<TB_, float, double>), tiles_grid(b), dim3(TB_), 0, s, b, src, dst_d));
  else INSP_TB(hipLaunchKernelGGL((kb_convert<TB_, double, double>), tiles_grid(b), dim3(TB_), 0, s, b, src, dst_d));
  check("kb_convert");
}

template <typename T>
void cheb(const BatchView& b, int level, int step, bool zero_init, hipStream_t s) {
  INSP_TB(hipLaunchKernelGGL((kb_cheb<TB_, T>), tiles_grid(b), dim3(TB_), 0, s, b, level, step, (int)zero_init));
  check("kb_cheb");
}

template <typename T>
void sub_from_b(const BatchView& b, int level, hipStream_t s) {
  INSP_TB(hipLaunchKernelGGL((kb_sub_from_b<TB_, T>), tiles_grid(b), dim3(TB_), 0, s, b, level));
  check("kb_sub_from_b");
}

template <typename T>
void vertex_restrict(const BatchView& b, hipStream_t s) {
  hipLaunchKernelGGL((kb_vertex_restrict<T>), dim3((unsigned)b.ltile_cap), dim3(kB), 0, s, b);
  check("kb_vertex_restrict");
}

template <typename T>
void vertex_prolong(const BatchView& b, hipStream_t s) {
  INSP_TB(hipLaunchKernelGGL((kb_vertex_prolong<TB_, T>), tiles_grid(b), dim3(TB_), 0, s, b));
  check("kb_vertex_prolong");
}

template <typename T>
void amg_spmv(const BatchView& b, int k, bool residual, hipStream_t s) {
  hipLaunchKernelGGL((kb_amg_spmv<T>), dim3((unsigned)b.ltile_cap), dim3(kB), 0, s, b, k, (int)residual);
  check("kb_amg_spmv");
}

template <typename T>
void amg_cheb(const BatchView& b, int k, int step, bool zero_init, hipStream_t s) {
  hipLaunchKernelGGL((kb_amg_cheb<T>), dim3((unsigned)b.ltile_cap), dim3(kB), 0, s, b, k, step,
                     (int)zero_init);
  check("kb_amg_cheb");
}

template <typename T>
void amg_restrict(const BatchView& b, int k, hipStream_t s) {
  hipLaunchKernelGGL((kb_amg_restrict<T>), dim3((unsigned)b.ltile_cap), dim3(kB), 0, s, b, k);
  check("kb_amg_restrict");
}

template <typename T>
void amg_prolong(const BatchView& b, int k, hipStream_t s) {
  hipLaunchKernelGGL((kb_amg_prolong<T>), dim3((unsigned)b.ltile_cap), dim3(kB), 0, s, b, k);
  check("kb_amg_prolong");
}

template <typename T>
void amg_dense(const BatchView& b, int k, hipStream_t s) {
  hipLaunchKernelGGL((kb_amg_dense<T>), dim3((unsigned)b.prob_cap), dim3(kB), 0, s, b, k);
  check("kb_amg_dense");
}

template <typename T>
void restrict_(const BatchView& b, const BatchTables& t, hipStream_t s) {
  INSP_DIMT(INSP_TB(hipLaunchKernelGGL((kb_restrict<TB_, D_, T>), tiles_grid(b), dim3(TB_), 0, s, b, t)));
  check("kb_restrict");
}

template <typename T>
void prolong(const BatchView& b, const BatchTables& t, hipStream_t s) {
  INSP_DIMT(INSP_TB(hipLaunchKernelGGL((kb_prolong<TB_, D_, T>), tiles_grid(b), dim3(TB_), 0, s, b, t)));
  check("kb_prolong");
}

template <typename T>
void cg_init(const BatchView& b, hipStream_t s) {
  INSP_TB(hipLaunchKernelGGL((kb_cg_init<TB_, T>), tiles_grid(b), dim3(TB_), 0, s, b));
  check("kb_cg_init");
}

template <typename T>
void cg_xr(const BatchView& b, int num, hipStream_t s) {
  INSP_TB(hipLaunchKernelGGL((kb_cg_xr<TB_, T>), tiles_grid(b), dim3(TB_), 0, s, b, num));
  check("kb_cg_xr");
}

template <typename T>
void cg_p(const BatchView& b, int num, int den, hipStream_t s) {
  INSP_TB(hipLaunchKernelGGL((kb_cg_p<TB_, T>), tiles_grid(b), dim3(TB_), 0, s, b, num, den));
  check("kb_cg_p");
}

void sum_d(const BatchView& b, int field, hipStream_t s) {
  INSP_TB(hipLaunchKernelGGL((kb_sum_d<TB_>), tiles_grid(b), dim3(TB_), 0, s, b, field));
  check("kb_sum_d");
}

void sub_mean_d(const BatchView& b, int field, hipStream_t s) {
  INSP_TB(hipLaunchKernelGGL((kb_sub_mean_d<TB_>), tiles_grid(b), dim3(TB_), 0, s, b, field));
  check("kb_sub_mean_d");
}

void rand_d(const BatchView& b, int level, int field, hipStream_t s) {
  INSP_TB(hipLaunchKernelGGL((kb_rand_d<TB_>), tiles_grid(b), dim3(TB_), 0, s, b, level, field));
  check("kb_rand_d");
}

void inv_apply_d(const BatchView& b, int level, int field, hipStream_t s) {
  INSP_TB(hipLaunchKernelGGL((kb_inv_apply_d<TB_>), tiles_grid(b), dim3(TB_), 0, s, b, level, field));
  check("kb_inv_apply_d");
}

void scale_d(const BatchView& b, int field, int slot, hipStream_t s) {
  INSP_TB(hipLaunchKernelGGL((kb_scale_d<TB_>), tiles_grid(b), dim3(TB_), 0, s, b, field, slot));
  check("kb_scale_d");
}

#define INSP_BINST(T)                                                                             \
  template void trace<T>(const BatchView&, const BatchTables&, int, int, hipStream_t);            \
  template void apply<T>(const BatchView&, const BatchTables&, int, int, int, int, bool,          \
                         hipStream_t);                                                            \
  template void dot<T>(const BatchView&, int, int, int, int, int, int, hipStream_t);              \
  template void cheb<T>(const BatchView&, int, int, bool, hipStream_t);                           \
  template void sub_from_b<T>(const BatchView&, int, hipStream_t);                                \
  template void vertex_restrict<T>(const BatchView&, hipStream_t);                                \
  template void vertex_prolong<T>(const BatchView&, hipStream_t);                                 \
  template void amg_spmv<T>(const BatchView&, int, bool, hipStream_t);                            \
  template void amg_cheb<T>(const BatchView&, int, int, bool, hipStream_t);                       \
  template void amg_restrict<T>(const BatchView&, int, hipStream_t);                              \
  template void amg_prolong<T>(const BatchView&, int, hipStream_t);                               \
  template void amg_dense<T>(const BatchView&, int, hipStream_t);                                 \
  template void restrict_<T>(const BatchView&, const BatchTables&, hipStream_t);                  \
  template void prolong<T>(const BatchView&, const BatchTables&, hipStream_t);                    \
  template void cg_init<T>(const BatchView&, hipStream_t);                                        \
  template void cg_xr<T>(const BatchView&, int, hipStream_t);                                     \
  template void cg_p<T>(const BatchView&, int, int, hipStream_t);
INSP_BINST(double)
INSP_BINST(float)

}  // namespace batch
}  // namespace insp
