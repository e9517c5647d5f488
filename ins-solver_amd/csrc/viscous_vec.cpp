#include "viscous_vec.h"

#include <chrono>
#include <cmath>

#include "kernels.h"
#include "refelem.h"
#include "vec_kernels.h"

namespace insp {

namespace {
enum VecSlot { V_RZ = 0, V_PAP = 1, V_RR = 2, V_RZN = 3, V_BNUM = 4, V_BB = 5, V_COUNT = 8 };
}  // namespace

VecViscousSolver::VecViscousSolver(HostMesh mesh, int order) : base_(std::move(mesh), order) {
  // the scalar solver only supplies the operator: one level, no smoother set-up
  base_.set_solver(SolverType::JACOBI_PCG);
  const HostMesh& m = base_.mesh_;
  ncomp_ = m.dim;
  const int64_t B = m.n_bfaces;
  types_.assign((size_t)ncomp_ * B, 0);
  // outward unit normal of every boundary face: n ~ sum_r alpha_fr grad(xi_r)
  const int d = m.dim;
  std::vector<double> G((size_t)m.n_cells * d * d);
  mesh_inverse_jacobians(m, G.data());
  std::vector<double> bn((size_t)std::max<int64_t>(B, 1) * d, 0.0);
  for (int64_t b = 0; b < B; ++b) {
    const int64_t e = m.bface2cells[b];
    const int f = m.bfaceNum[b];
    double n[3] = {0, 0, 0}, nn = 0;
    for (int x = 0; x < d; ++x) {
      for (int r = 0; r < d; ++r) n[x] += face_alpha(d, f, r) * G[((size_t)e * d + r) * d + x];
      nn += n[x] * n[x];
    }
    nn = std::sqrt(nn);
    for (int x = 0; x < d; ++x) bn[(size_t)b * d + x] = n[x] / nn;
  }
  bn_.upload(bn.data(), bn.size());
  std::vector<int> eff((size_t)ncomp_ * std::max<int64_t>(B, 1), 0);
  eff_.upload(eff.data(), eff.size());
  scal_.alloc(V_COUNT);
  INSP_HIP(hipHostMalloc((void**)&host_scal_, V_COUNT * sizeof(double), hipHostMallocDefault));
}

VecViscousSolver::~VecViscousSolver() {
  (void)hipStreamSynchronize(base_.stream_);
  if (host_scal_) (void)hipHostFree(host_scal_);
}

void VecViscousSolver::set_bc_types(const int* const types[], int64_t n) {
  const int64_t B = base_.mesh_.n_bfaces;
  if (n != B) throw InvalidArgument("bc_types length must equal the number of bfaces");
  if (B == 0) return;
  for (int c = 0; c < ncomp_; ++c)
    if (!types || !types[c]) throw InvalidArgument("bc_types of every component are required");
  std::vector<int> eff((size_t)ncomp_ * B);
  for (int64_t b = 0; b < B; ++b) {
    bool slip = true;
    for (int c = 0; c < ncomp_; ++c) {
      const int t = types[c][b];
      if (t < 0 || t > 2) throw InvalidArgument("bc types are 0 (Dirichlet), 1 (Neumann) or 2 (slip)");
      types_[(size_t)c * B + b] = t;
      slip = slip && t == 2;
    }
    for (int c = 0; c < ncomp_; ++c) {
      const int t = types[c][b];
      eff[(size_t)c * B + b] = t == 2 ? (slip ? 2 : 3) : t;
    }
  }
  eff_.upload(eff.data(), eff.size());
  initialised_ = false;
}

void VecViscousSolver::set_bcs(const double* const bc[], Layout layout) {
  const int64_t B = base_.mesh_.n_bfaces;
  const int Nfp = base_.nfp();
  if (!bc) {
    have_bc_ = false;
    bc_.release();
    return;
  }
  for (int c = 0; c < ncomp_; ++c)
    if (!bc[c]) throw InvalidArgument("bc data of every component are required");
  if (B == 0) return;
  const size_t per = (size_t)B * Nfp;
  bc_.alloc((size_t)ncomp_ * per, false);
  for (int c = 0; c < ncomp_; ++c) {
    const hipMemcpyKind kind = layout == Layout::HOST_AOS ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    INSP_HIP(hipMemcpyAsync(bc_.p + c * per, bc[c], per * sizeof(double), kind, base_.stream_));
  }
  INSP_HIP(hipStreamSynchronize(base_.stream_));
  have_bc_ = true;
}

void VecViscousSolver::set_tol_and_iter(double rtol, double atol, int max_iter) {
  if (!(rtol >= 0) || !(atol >= 0) || max_iter < 0) throw InvalidArgument("invalid tolerances");
  rtol_ = rtol;
  atol_ = atol;
  max_iter_ = max_iter;
}

void VecViscousSolver::init() {
  if (!base_.initialised_) base_.init();
  const int64_t E = base_.mesh_.n_cells, st = base_.stride_;
  const int Np = base_.np();
  const size_t n = (size_t)ncomp_ * Np * st;
  b_.alloc(n); x_.alloc(n); r_.alloc(n); z_.alloc(n); zold_.alloc(n); p_.alloc(n); Ap_.alloc(n);
  tr_.alloc((size_t)ncomp_ * base_.fine_face_nodes() * st);
  partial_.alloc((size_t)std::max<int64_t>(4 * kRedBlocks, (int64_t)ncomp_ * apply_blocks(E)));
  const std::vector<double> Mi = mat_inverse(base_.fine_re_.M, Np);
  minv_.upload(Mi.data(), Mi.size());
  inv_.release();
  blk_.release();
  // exact diagonal of every component (the Jacobi preconditioner's, and ins_viscous_vec_diag's)
  diag_.alloc(n);
  {
    const MeshView mv0 = base_.mesh_view();
    const LevelView lv = base_.level_view(0, false);
    const double* k = base_.fine_factor();
    for (int c = 0; c < ncomp_; ++c) {
      MeshView mv = mv0;
      mv.bctype = eff_.p + (size_t)c * std::max<int64_t>(base_.mesh_.n_bfaces, 1);
      double* dc = diag_.p + (size_t)c * Np * st;
      launch_diag(mv, lv, k, dc, base_.stream_, bn_.p, c);
      if (base_.has_mass_term())
        launch_mass_diag(mv, lv, Np, base_.fine_mm(), base_.mm_const_, dc,
                         base_.stream_);
    }
  }
  if (prec_ == VecPrecond::JACOBI) {
    inv_.alloc(n);
    reciprocal<double>(ncomp_ * Np, E, st, diag_.p, inv_.p, base_.stream_);
  }
  INSP_HIP(hipStreamSynchronize(base_.stream_));
  if (prec_ == VecPrecond::BLOCK_JACOBI) build_block_jacobi();
  initialised_ = true;
}

// Traces of every component (orders >= 2), then the element pass of each component.
void VecViscousSolver::apply(const double* u, double* out, double* partial) {
  const MeshView mv = base_.mesh_view();
  const LevelView lv = base_.level_view(0, false);
  const double* k = base_.fine_factor();
  const int64_t cs = (int64_t)base_.np() * base_.stride_;
  const int64_t ts = (int64_t)base_.fine_face_nodes() * base_.stride_;
  for (int c = 0; c < ncomp_; ++c) launch_trace<double>(mv, lv, u + c * cs, k, tr_.p + c * ts, base_.stream_);
  for (int c = 0; c < ncomp_; ++c) apply_one(c, u, out + c * cs, partial);
}

// Component c of A u (the traces of every component already in tr_).
void VecViscousSolver::apply_one(int c, const double* u, double* out_c, double* partial) {
  MeshView mv = base_.mesh_view();
  mv.bctype = eff_.p + (size_t)c * std::max<int64_t>(base_.mesh_.n_bfaces, 1);
  const LevelView lv = base_.level_view(0, false);
  const double* k = base_.fine_factor();
  const int Np = base_.np();
  const int64_t cs = (int64_t)Np * base_.stride_;
  const int64_t ts = (int64_t)base_.fine_face_nodes() * base_.stride_;
  double* pc = partial ? partial + (size_t)c * apply_blocks(base_.mesh_.n_cells) : nullptr;
  const SlipFaces<double> slip{u, tr_.p, cs, ts, bn_.p, c};
  launch_apply_slip(mv, lv, u + c * cs, tr_.p + c * ts, k, out_c, slip, base_.stream_, pc);
  if (base_.has_mass_term())
    launch_mass_add<double>(mv, lv, Np, base_.fine_mm(),
                            base_.mm_const_, u + c * cs, out_c, base_.stream_, pc);
}

void VecViscousSolver::build_block_jacobi() {
  // the (c, c) block of every cell, probed as the scalar solver probes its blocks: a unit vector
  // on every cell of one colour of the face graph returns each of those cells' own columns
  const HostMesh& m = base_.mesh_;
  const int64_t E = m.n_cells, st = base_.stride_;
  const int Nf = m.dim + 1, Np = base_.np();
  std::vector<int> colour((size_t)E, -1);
  int ncol = 0;
  for (int64_t e = 0; e < E; ++e) {
    unsigned used = 0;
    for (int f = 0; f < Nf; ++f) {
      const int n = m.nbr[(size_t)e * Nf + f];
      if (n >= 0 && colour[n] >= 0) used |= 1u << colour[n];
    }
    int c = 0;
    while (used & (1u << c)) ++c;
    colour[e] = c;
    ncol = std::max(ncol, c + 1);
  }
  DBuf<int> dcol;
  dcol.upload(colour.data(), colour.size());
  const size_t cs = (size_t)Np * st, bs = (size_t)Np * Np * st;
  blk_.alloc((size_t)ncomp_ * bs);
  DBuf<double> u, y;
  u.alloc((size_t)ncomp_ * cs);  // zero outside the probed component
  y.alloc(cs);
  fill<double>((int64_t)tr_.n, tr_.p, 0.0, base_.stream_);  // zero traces of the other components
  const MeshView mv = base_.mesh_view();
  const LevelView lv = base_.level_view(0, false);
  const double* k = base_.fine_factor();
  const int64_t ts = (int64_t)base_.fine_face_nodes() * st;
  for (int comp = 0; comp < ncomp_; ++comp) {
    for (int c = 0; c < ncol; ++c)
      for (int j = 0; j < Np; ++j) {
        block_probe_unit(Np, E, st, dcol.p, c, j, u.p + comp * cs, base_.stream_);
        launch_trace<double>(mv, lv, u.p + comp * cs, k, tr_.p + comp * ts, base_.stream_);
        apply_one(comp, u.p, y.p, nullptr);
        block_extract_column(Np, E, st, dcol.p, c, j, y.p, blk_.p + comp * bs, base_.stream_);
      }
    fill<double>((int64_t)cs, u.p + comp * cs, 0.0, base_.stream_);
    fill<double>(ts, tr_.p + comp * ts, 0.0, base_.stream_);
    block_invert(Np, E, st, blk_.p + comp * bs, base_.stream_);
  }
  INSP_HIP(hipStreamSynchronize(base_.stream_));
}

void VecViscousSolver::inv_block_diag(double* out) {
  if (!initialised_) init();
  if (prec_ != VecPrecond::BLOCK_JACOBI || blk_.empty())
    throw InvalidArgument("inverse blocks exist only with the block-Jacobi preconditioner");
  const int Np = base_.np();
  const int64_t E = base_.mesh_.n_cells, st = base_.stride_;
  std::vector<double> h((size_t)ncomp_ * Np * Np * st);
  INSP_HIP(hipMemcpyAsync(h.data(), blk_.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, base_.stream_));
  INSP_HIP(hipStreamSynchronize(base_.stream_));
  for (int c = 0; c < ncomp_; ++c)
    for (int64_t e = 0; e < E; ++e)
      for (int k = 0; k < Np * Np; ++k)
        out[((size_t)c * E + e) * Np * Np + k] = h[(size_t)c * Np * Np * st + (size_t)k * st + e];
}

void VecViscousSolver::precondition(const double* r, double* z) {
  const int Np = base_.np();
  const int64_t E = base_.mesh_.n_cells, st = base_.stride_;
  const int64_t cs = (int64_t)Np * st, n = ncomp_ * cs;
  hipStream_t s = base_.stream_;
  switch (prec_) {
    case VecPrecond::NONE:
      copy<double>(n, r, z, s);
      break;
    case VecPrecond::INV_MASS:
    case VecPrecond::FACTOR_INV_MASS:
    case VecPrecond::RECP_FACTOR_DAT_INV_MASS: {
      // pre_inv_mass / pre_factor_inv_mass / pre_recp_dat_factor_inv_mass
      // (src/solvers/2d/slip_matrix/viscous_solver.cpp:334-366)
      const double scale = prec_ == VecPrecond::FACTOR_INV_MASS ? base_.inv_mass_factor_ : 1.0;
      const double* recp = nullptr;
      if (prec_ == VecPrecond::RECP_FACTOR_DAT_INV_MASS) {
        if (base_.inv_mass_recp_.empty())
          throw InvalidArgument("the reciprocal-factor inverse-mass preconditioner needs set_inv_mass_recp_factor");
        recp = base_.inv_mass_recp_.p;
      }
      for (int c = 0; c < ncomp_; ++c)
        launch_inv_mass(Np, E, st, base_.J_.p, minv_.p, scale, recp, r + c * cs, z + c * cs, s);
      break;
    }
    case VecPrecond::JACOBI:
      jacobi<double>(n, r, inv_.p, z, s);
      break;
    case VecPrecond::BLOCK_JACOBI:
      for (int c = 0; c < ncomp_; ++c)
        block_apply(Np, E, st, blk_.p + (size_t)c * Np * Np * st, r + c * cs, z + c * cs, s);
      break;
  }
}

void VecViscousSolver::import(const double* const src[], double* dst, Layout layout) {
  if (!src) throw InvalidArgument("null vector");
  const int64_t cs = (int64_t)base_.np() * base_.stride_;
  for (int c = 0; c < ncomp_; ++c) base_.import_vec(src[c], dst + c * cs, layout, base_.np());
}

void VecViscousSolver::export_to(const double* src, double* const dst[], Layout layout) {
  if (!dst) throw InvalidArgument("null vector");
  const int64_t cs = (int64_t)base_.np() * base_.stride_;
  for (int c = 0; c < ncomp_; ++c) base_.export_vec(src + c * cs, dst[c], layout, base_.np());
}

double VecViscousSolver::read_scalars() {
  INSP_HIP(hipMemcpyAsync(host_scal_, scal_.p, V_COUNT * sizeof(double), hipMemcpyDeviceToHost, base_.stream_));
  INSP_HIP(hipStreamSynchronize(base_.stream_));
  return host_scal_[0];
}

void VecViscousSolver::mult(const double* const in[], double* const out[], Layout layout) {
  if (!initialised_) init();
  import(in, b_.p, layout);
  apply(b_.p, Ap_.p, nullptr);
  export_to(Ap_.p, out, layout);
  INSP_HIP(hipStreamSynchronize(base_.stream_));
}

void VecViscousSolver::diag(double* const out[], Layout layout) {
  if (!initialised_) init();
  export_to(diag_.p, out, layout);
  INSP_HIP(hipStreamSynchronize(base_.stream_));
}

// ViscousMatrix2D::apply_bc (src/solvers/2d/slip_matrix/viscous_matrix.cpp:26-43): the terms of
// each component with its own bc type; slip faces add nothing.
void VecViscousSolver::apply_bc(double* const rhs[], Layout layout) {
  if (!initialised_) init();
  import(rhs, b_.p, layout);
  const int64_t cs = (int64_t)base_.np() * base_.stride_;
  const int64_t B = base_.mesh_.n_bfaces;
  const LevelView lv = base_.level_view(0, false);
  const double* k = base_.fine_factor();
  if (have_bc_)
    for (int c = 0; c < ncomp_; ++c) {
      MeshView mv = base_.mesh_view();
      mv.bctype = eff_.p + (size_t)c * std::max<int64_t>(B, 1);
      launch_apply_bc(mv, lv, bc_.p + (size_t)c * B * base_.nfp(), k, b_.p + c * cs, base_.stream_);
    }
  export_to(b_.p, rhs, layout);
  INSP_HIP(hipStreamSynchronize(base_.stream_));
}

// ViscousSolver2D::solve (src/solvers/2d/slip_matrix/viscous_solver.cpp:59-75): boundary terms
// into the rhs, then (preconditioned) CG over all components at once; inner products sum over
// the components. As the scalar solve: flexible CG, true relative residual as the stopping test.
int VecViscousSolver::solve(const double* const rhs[], double* const ans[], Layout layout, SolveStats* stats) {
  if (!initialised_) init();
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = base_.stream_;
  const int Np = base_.np();
  const int64_t E = base_.mesh_.n_cells, st = base_.stride_;
  const int64_t cs = (int64_t)Np * st, n = ncomp_ * cs;
  const int64_t B = base_.mesh_.n_bfaces;
  RedScratch rs{partial_.p};
  import(rhs, b_.p, layout);
  if (have_bc_) {
    const LevelView lv = base_.level_view(0, false);
    const double* k = base_.fine_factor();
    for (int c = 0; c < ncomp_; ++c) {
      MeshView mv = base_.mesh_view();
      mv.bctype = eff_.p + (size_t)c * std::max<int64_t>(B, 1);
      launch_apply_bc(mv, lv, bc_.p + (size_t)c * B * base_.nfp(), k, b_.p + c * cs, s);
    }
  }
  import(const_cast<const double* const*>(ans), x_.p, layout);
  dot<double>(n, b_.p, b_.p, nullptr, nullptr, 1, scal_.p, V_BB, rs, s);
  read_scalars();
  const double bnorm = std::sqrt(host_scal_[V_BB]);
  int status = 1, it = 0;
  double rr_true = 0;
  int64_t applies = 0;
  const int nblk = ncomp_ * apply_blocks(E);
  if (!std::isfinite(bnorm)) {
    status = -3;
  } else {
    const double tol = std::max(rtol_ * bnorm, atol_);
    for (int restart = 0;; ++restart) {
      apply(x_.p, Ap_.p, nullptr);
      ++applies;
      residual(n, b_.p, Ap_.p, r_.p, scal_.p, V_RR, rs, s);
      read_scalars();
      rr_true = host_scal_[V_RR];
      if (!std::isfinite(rr_true)) { status = -3; break; }
      if (std::sqrt(rr_true) <= tol) { status = 0; break; }
      if (it >= max_iter_ || restart > 50) { status = 1; break; }
      precondition(r_.p, z_.p);
      dot<double>(n, r_.p, z_.p, nullptr, nullptr, 1, scal_.p, V_RZ, rs, s);
      copy<double>(n, z_.p, p_.p, s);
      bool breakdown = false;
      while (it < max_iter_) {
        apply(p_.p, Ap_.p, partial_.p);  // p.Ap fused into the element passes
        ++applies;
        finalize(nblk, 1, partial_.p, scal_.p, V_PAP, s);
        cg_update_xr<double>(n, x_.p, r_.p, p_.p, Ap_.p, nullptr, nullptr, scal_.p, V_RZ, V_PAP, V_RR, rs, s);
        ++it;
        read_scalars();
        const double pap = host_scal_[V_PAP], rr = host_scal_[V_RR];
        if (!(pap > 0) || !std::isfinite(rr) || !std::isfinite(pap)) { breakdown = true; break; }
        if (std::sqrt(rr) <= tol) break;
        std::swap(z_.p, zold_.p);
        precondition(r_.p, z_.p);
        dot<double>(n, r_.p, z_.p, nullptr, nullptr, 1, scal_.p, V_RZN, rs, s);
        dot_flex<double>(n, r_.p, z_.p, zold_.p, scal_.p, V_BNUM, rs, s);
        cg_update_p<double>(n, p_.p, z_.p, scal_.p, V_BNUM, V_RZ, s);
        scal_copy(scal_.p, V_RZ, V_RZN, s);
      }
      if (breakdown) { status = -3; break; }
    }
  }
  export_to(x_.p, ans, layout);
  INSP_HIP(hipStreamSynchronize(s));
  const auto t1 = std::chrono::steady_clock::now();
  SolveStats stt;
  stt.bnorm = bnorm;
  stt.iterations = it;
  stt.converged = status == 0;
  stt.abs_residual = std::sqrt(rr_true);
  stt.rel_residual = bnorm > 0 ? stt.abs_residual / bnorm : 0.0;
  stt.seconds = std::chrono::duration<double>(t1 - t0).count();
  stt.fine_applies = applies;
  if (stats) *stats = stt;
  return status;
}

}  // namespace insp
