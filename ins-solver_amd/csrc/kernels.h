// Host-side launchers for the HIP kernels. All vectors use the OP2 GPU
// struct-of-arrays layout: value i of cell e lives at v[i*stride + e]; entries
// of padding cells (e >= E) are kept at zero.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace insp {

struct MeshView {
  int dim;
  int64_t E, stride;
  const double* Kt;    // [NK][stride] metric scaled by J
  const double* J;     // [stride]
  const int* nbr;      // [Nf][stride] neighbour or -(bface+1)
  const int* code;     // [Nf][stride] neighbour face * nperm + vertex permutation
  const int* bctype;   // [n_bfaces] 0 Dirichlet, 1 Neumann
  // [Nf][stride] (neighbour << 5) | code inside the mesh, -(bface+1) on the boundary; null when
  // the mesh has 2^26 cells or more (the element passes then read nbr and code)
  const int* nbc = nullptr;
  // [E][NK] the metric cell by cell (3D; null otherwise): the order-1 pass gathers a
  // neighbour's metric as three 16-byte loads from one or two cache lines instead of six rows
  const double* Ka = nullptr;
  // [E][4] the packed words cell by cell (3D, when packed; null otherwise), for the coarse CG's
  // order-1 apply
  const int* nbc_cells = nullptr;
};

// Boundary faces of the vector (velocity) operator with the slip condition (bc type 2 on
// every component: vmf_2d_mult_avg_jump.h:7-27, vmf_3d_mult_avg_jump.h:8-40): component c
// sees the Dirichlet ghost of its share of the normal part, u_c -> n_c (n.U) and
// T_c -> n_c sum_d n_d T_d, so the normal velocity is held at zero and the tangential part
// is free. The components of U live cstride apart, their face traces tstride apart.
struct NoSlip {
  static constexpr bool enabled = false;
};
template <typename T>
struct SlipFaces {
  static constexpr bool enabled = true;
  const T* u0;        // component 0 of the vector
  const T* tr0;       // face traces of component 0 (orders >= 2)
  int64_t cstride, tstride;
  const double* bn;   // [n_bfaces][DIM] outward unit normals
  int comp;
};

struct LevelView {
  int N;
  const void* C;       // operator constants in the kernel precision
  const double* C64;   // operator constants in double (set-up kernels)
  const int* tab;
  const double* tauS;  // [Nf][stride] penalty * sJ
  const void* frag = nullptr;  // matrix-core operator fragments in the kernel precision, or null
  const double* tau_cells = nullptr;  // [E][4] tauS cell by cell (3D order 1), or null
};

// SIPG operator: trace pass + element pass (order 1: element pass only). With
// `partial` the element pass also writes per-block partial sums of u.(A u)
// (apply_blocks(E) of them) for a deterministic finalize.
template <typename T>
void launch_trace(const MeshView& m, const LevelView& l, const T* u, const T* kfac, T* tr,
                  hipStream_t s);
template <typename T>
void launch_apply(const MeshView& m, const LevelView& l, const T* u, const T* tr, const T* kfac,
                  T* out, hipStream_t s, double* partial = nullptr);
int apply_blocks(int64_t E);
// Order-1 apply (3D, with the cell-by-cell metric) on values held cell by cell: u and out are
// [E][4] (the coarse Jacobi-CG runs on such copies); the coefficient stays in rows.
template <typename T>
void launch_apply_p1_cells(const MeshView& m, const LevelView& l, const T* u, const T* kfac, T* out,
                           hipStream_t s, double* partial = nullptr);
// Element pass (vector kernels, order 1 included) with one Chebyshev step of the smoother in
// place of the store, in three-term form: res = inv (f - A u), x_next = u + c1 (u - prev) +
// c2 res (prev not read when c1 = 0, zero when null). x_next must differ from u and may be prev.
// Returns false, launching nothing, where the level takes the matrix-core element pass.
template <typename T>
bool launch_apply_cheb(const MeshView& m, const LevelView& l, const T* u, const T* tr,
                       const T* kfac, const T* f, const T* inv, const T* prev, T* x_next, T c1,
                       T c2, hipStream_t s);
// Element pass of component slip.comp of the vector (velocity) operator, FP64: the scalar
// operator of that component, except that faces of bc type 2 are slip faces, which read the
// other components' values and face traces of the cell (the traces of every component from
// launch_trace first, at orders >= 2). With `partial`, u.(A u) partials as launch_apply.
void launch_apply_slip(const MeshView& m, const LevelView& l, const double* u, const double* tr,
                       const double* kfac, double* out, const SlipFaces<double>& slip, hipStream_t s,
                       double* partial = nullptr);
// Over-integrated variable-coefficient operator (sipg_oi.hip; rules and tables in oi.h): the
// coefficient at the volume cubature points (kv, [nv][stride]) and the face points (ks,
// [Nf*ns][stride]); the trace pass writes u and sJ k n.grad u at every face point (tU, tT).
template <typename T>
struct OiView {
  const T* C;
  const int* qmap;
  const T* kv;
  const T* ks;
  T* tU;
  T* tT;
};
template <typename T>
void launch_apply_oi(const MeshView& m, const LevelView& l, const OiView<T>& o, const T* u, T* out,
                     hipStream_t s, double* partial = nullptr);
void launch_diag_oi(const MeshView& m, const LevelView& l, const OiView<double>& o, double* diag,
                    hipStream_t s);
// The kernels' compile-time rule sizes agree with the host's rule and tables.
bool oi_shape_matches(int dim, int N, int nv, int ns, int table_size);
// Element pass on the matrix cores (sipg_mfma.hip): 3D, orders 2-4. launch_apply
// takes it whenever the level view carries fragments.
bool mfma_apply_supported(int dim, int N);
// A-operand fragments of one order built from the operator constants and index
// table (make_op_tables); empty when the order has no matrix-core pass.
std::vector<double> build_mfma_fragments(int dim, int N, const std::vector<double>& C,
                                         const std::vector<int>& tab);
template <typename T>
void launch_apply_mfma(const MeshView& m, const LevelView& l, const T* u, const T* tr,
                       const T* kfac, T* out, hipStream_t s, double* partial);
// Exact diagonal of the operator (double). With bnormal ([n_bfaces][dim] outward unit
// normals): the diagonal of component comp of the vector operator, whose bc type 2 faces are
// slip faces.
void launch_diag(const MeshView& m, const LevelView& l, const double* kfac, double* diag,
                 hipStream_t s, const double* bnormal = nullptr, int comp = 0);
// Penalty tau*sJ per cell face (double), tau = 2(P+1)(P+2) max(fscale k) with P = pen_order.
void launch_tau(const MeshView& m, const LevelView& l, const double* kfac, int pen_order,
                double* tauS, hipStream_t s);
// rhs += boundary-condition terms; bc is bfaces x Nfp (row per boundary face).
void launch_apply_bc(const MeshView& m, const LevelView& l, const double* bc, const double* kfac,
                     double* rhs, hipStream_t s);
// out = J * M * in per cell.
void launch_mass(const MeshView& m, const LevelView& l, const double* in, double* out,
                 hipStream_t s);
// y (+)= A x per cell, A is rows x cols row-major (device, precision T).
template <typename T>
void launch_elem_matvec(int dim, int rows_order, int cols_order, int64_t E, int64_t stride,
                        const T* A, const T* x, T* y, bool add, hipStream_t s);

// Block-Jacobi (block_kernels.hip): blocks SoA over cells, (i, j) of cell e at
// blk[(i*rows + j)*stride + e].
void block_probe_unit(int rows, int64_t E, int64_t stride, const int* colour, int c, int j, double* u,
                      hipStream_t s);
void block_extract_column(int rows, int64_t E, int64_t stride, const int* colour, int c, int j,
                          const double* y, double* blk, hipStream_t s);
void block_invert(int rows, int64_t E, int64_t stride, double* blk, hipStream_t s);
void block_apply(int rows, int64_t E, int64_t stride, const double* blk, const double* r, double* z,
                 hipStream_t s);

// Mass term of the viscous (Helmholtz) operator (mass_kernels.hip): out += J M (mm o u),
// or mmc J M u when mm is null; adds its share of u.(A u) to `partial` (same blocks as
// launch_apply). np is the level's node count.
template <typename T>
void launch_mass_add(const MeshView& m, const LevelView& l, int np, const T* mm, T mmc, const T* u,
                     T* out, hipStream_t s, double* partial);
// diag += mm_i J M_ii (or mmc J M_ii).
void launch_mass_diag(const MeshView& m, const LevelView& l, int np, const double* mm, double mmc,
                      double* diag, hipStream_t s);
// Inverse-mass preconditioner: z = scale (J M)^-1 r, or ((J M)^-1 r)_i / recp_i when recp
// is not null; Minv is the reference element's inverse mass matrix (np x np, row-major).
void launch_inv_mass(int np, int64_t E, int64_t stride, const double* J, const double* Minv,
                     double scale, const double* recp, const double* r, double* z, hipStream_t s);

// Gradient with central flux (proj_kernels.hip): out[x] = grad_h p, or u[x] - dt grad_h p
// when u[x] is not null (x < dim; unused entries null). G: grad(xi) per cell, SoA rows
// (r*dim + x); lift: Np x (Nf*Nfp) = M^-1 EMAT of the order.
void launch_grad_central(int dim, int N, int64_t E, int64_t stride, const double* G, const int* nbr,
                         const int* code, const double* C64, const int* tab, const double* lift,
                         const double* p, double dt, const double* const u[3], double* const out[3],
                         hipStream_t s);

// Scratch of the allocation-free apply (sipg_kernels.hip): a static device array of
// kNoAllocScratchBytes per GPU. Returns the current device's copy, looked up once per device.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
double* noalloc_scratch();

// Achievable peaks of the current device (probe_kernels.hip): streaming copy bandwidth
// (GB/s, read + write) and FP32 / FP64 16x16x4 matrix-core throughput (TFLOP/s).
void measure_device_peaks(double* copy_gbs, double* mfma_f32_tflops, double* mfma_f64_tflops);

// Pressure rhs of the caller (rhs_kernels.hip): div <- J M (-div/dt) + J EMAT g with
// g = b0 dPdN + b1 dPdNOld written back to dPdNOld; M is Np x Np, EMAT Np x (Nf*Nfp).
void launch_pressure_rhs(int dim, int N, int64_t E, int64_t stride, const double* J, const double* M,
                         const double* EMAT, double b0, double b1, double dt, const double* dPdN,
                         double* dPdNOld, double* div, hipStream_t s);

}  // namespace insp
