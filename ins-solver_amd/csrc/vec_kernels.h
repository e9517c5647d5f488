// Vector kernels of the Krylov solvers and smoothers. Reductions are
// two-stage and deterministic: a fixed grid (a function of the vector length
// only) writes per-block partial sums in double, and one block adds them in
// index order into a device scalar slot. Step lengths are computed on the
// device from those slots, so no host round trip is needed inside a cycle.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "reduce.h"

namespace insp {

struct RedScratch {
  double* partial;  // max(kRedBlocks, apply blocks) * 4 doubles
};

int red_blocks(int64_t n);

// scal[slot + k] = sum of the first nb partials of value k (partial[k * kRedBlocks + b]),
// added in index order by one block.
void finalize(int nb, int nvals, const double* partial, double* scal, int slot, hipStream_t s);
// scal[slot+k] = sum(a_k * b_k) for k < nd (nd = 1 or 2); c,d may be null for nd = 1.
template <typename T>
void dot(int64_t n, const T* a, const T* b, const T* c, const T* d, int nd, double* scal, int slot,
         RedScratch rs, hipStream_t s);
// scal[slot] = sum(r * (z - zold))
template <typename T>
void dot_flex(int64_t n, const T* r, const T* z, const T* zold, double* scal, int slot,
              RedScratch rs, hipStream_t s);
// scal[slot] = sum(r * z), scal[slot + 1] = sum(r * (z - zold)), in one pass (each sum in the
// order dot and dot_flex use)
template <typename T>
void dot_rz_flex(int64_t n, const T* r, const T* z, const T* zold, double* scal, int slot,
                 RedScratch rs, hipStream_t s);
// r = b - Ax, scal[slot] = |r|^2
void residual(int64_t n, const double* b, const double* Ax, double* r, double* scal, int slot,
              RedScratch rs, hipStream_t s);
// r = b - Ax computed in FP64 and stored in FP32; scal[slot] = |b - Ax|^2 (FP64)
void residual_f32(int64_t n, const double* b, const double* Ax, float* r, double* scal, int slot,
                  RedScratch rs, hipStream_t s);
// alpha = scal[num]/scal[den]; x (FP64) += alpha p; r (FP32) -= alpha Ap; scal[out] = |r|^2
// accumulated in FP64
void cg_update_xr_mixed(int64_t n, double* x, float* r, const float* p, const float* Ap, double* scal,
                        int num, int den, int out, RedScratch rs, hipStream_t s);
// alpha = scal[num]/scal[den]; x += alpha p; r -= alpha Ap; scal[out] = |r|^2
// (or r.z with z = invdiag*r when invdiag is given).
template <typename T>
void cg_update_xr(int64_t n, T* x, T* r, const T* p, const T* Ap, const T* invdiag, T* z,
                  double* scal, int num, int den, int out, RedScratch rs, hipStream_t s);
// beta = scal[num]/scal[den]; p = z + beta p
template <typename T>
void cg_update_p(int64_t n, T* p, const T* z, const double* scal, int num, int den, hipStream_t s);
// Jacobi-CG in z = invdiag * r: alpha = scal[num]/scal[den]; z -= alpha invdiag Ap;
// scal[out] = r.z = sum z^2 / invdiag (entries with invdiag = 0 add nothing)
template <typename T>
void cg_update_z(int64_t n, T* z, const T* Ap, const T* invdiag, double* scal, int num, int den,
                 int out, RedScratch rs, hipStream_t s);
// alpha = scal[anum]/scal[aden], beta = scal[bnum]/scal[anum]; x += alpha p; p = z + beta p
// unless last
template <typename T>
void cg_update_xp(int64_t n, T* x, T* p, const T* z, const double* scal, int anum, int aden,
                  int bnum, bool last, hipStream_t s);
// p_next = z + beta p (beta = scal[bnum]/scal[anum]); alpha = scal[anum]/scal[aden] is stored in
// scal[asave] for the x update of the next iteration. p_next must differ from p.
template <typename T>
void cg_update_p(int64_t n, T* p_next, const T* p, const T* z, double* scal, int anum, int aden,
                 int bnum, int asave, hipStream_t s);
// x += scal[asave] p0 + alpha p1 (alpha = scal[anum]/scal[aden]), then p0 = z + beta p1 unless last
// (beta = scal[bnum]/scal[anum]).
template <typename T>
void cg_update_x2p(int64_t n, T* x, T* p0, const T* p1, const T* z, const double* scal, int asave,
                   int anum, int aden, int bnum, bool last, hipStream_t s);
// z = invdiag * r
template <typename T>
void jacobi(int64_t n, const T* r, const T* invdiag, T* z, hipStream_t s);
// Chebyshev step: r = invdiag*(f - Ax) (Ax may be null = 0); d = c1*d + c2*r; x += d
// (x = d when x_zero). d may be null when c1 = 0 (then it is neither read nor written).
template <typename T>
void cheb_step(int64_t n, const T* f, const T* Ax, const T* invdiag, T* d, T* x, T c1, T c2,
               bool x_zero, hipStream_t s);
// Np values per cell from rows (v[i*stride + e]) to cell by cell (v[e*Np + i]) and back, out of
// place, over the whole Np*stride range (padding cells and the cell-by-cell tail written as zero).
template <typename T>
void rows_to_cells(int64_t E, int64_t stride, int Np, const T* src, T* dst, hipStream_t s);
template <typename T>
void cells_to_rows(int64_t E, int64_t stride, int Np, const T* src, T* dst, hipStream_t s);
// Chebyshev step in three-term form: x_next = x + c1 (x - prev) + c2 invdiag (f - Ax), prev not
// read when c1 = 0 and zero when null. x_next must differ from x and may be prev.
template <typename T>
void cheb_step3(int64_t n, const T* f, const T* Ax, const T* invdiag, const T* prev, const T* x,
                T* x_next, T c1, T c2, hipStream_t s);
// y = x converted
template <typename Tin, typename Tout>
void convert(int64_t n, const Tin* x, Tout* y, hipStream_t s);
template <typename T>
void fill(int64_t n, T* x, T v, hipStream_t s);
template <typename T>
void copy(int64_t n, const T* x, T* y, hipStream_t s);
// y = a*x + b*y
template <typename T>
void axpby(int64_t n, T a, const T* x, T b, T* y, hipStream_t s);
// x -= scal[slot]/count on valid cells (e < E) of an Np-row SoA vector.
template <typename T>
void sub_mean(int rows, int64_t E, int64_t stride, T* x, const double* scal, int slot,
              double count, hipStream_t s);
// Deterministic pseudo-random values in [-1,1) on valid cells, zero on padding.
void rand_fill(int rows, int64_t E, int64_t stride, uint64_t seed, double* x, hipStream_t s);
// scal[slot] = sum of x over valid cells.
template <typename T>
void vsum(int64_t n, const T* x, double* scal, int slot, RedScratch rs, hipStream_t s);
// SoA <-> AoS (device) transposition: aos[e*rows + i] <-> soa[i*stride + e].
void soa_from_aos(int rows, int64_t E, int64_t stride, const double* aos, double* soa,
                  hipStream_t s);
void aos_from_soa(int rows, int64_t E, int64_t stride, const double* soa, double* aos,
                  hipStream_t s);
// Zero the padding cells of an Np-row SoA vector.
template <typename T>
void zero_padding(int rows, int64_t E, int64_t stride, T* x, hipStream_t s);
// invdiag = 1/diag on valid cells (converted to T).
template <typename T>
void reciprocal(int rows, int64_t E, int64_t stride, const double* diag, T* inv, hipStream_t s);
// scal[dst] = scal[src]
void scal_copy(double* scal, int dst, int src, hipStream_t s);

}  // namespace insp
