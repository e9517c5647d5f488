// extern "C" boundary: exceptions never cross it; every entry point returns a
// status code and records a message for ins_poisson_last_error().
#include "../../include/ins_poisson.h"

#include <new>
#include <string>

#include "mesh.h"
#include "solver.h"

struct ins_mesh {
  insp::HostMesh m;
};
struct ins_poisson {
  std::unique_ptr<insp::PoissonSolver> s;
};

namespace {
thread_local std::string g_last_error;

template <typename F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const insp::InvalidArgument& e) {
    g_last_error = e.what();
    return INS_ERR_INVALID;
  } catch (const insp::HipError& e) {
    g_last_error = e.what();
    return INS_ERR_HIP;
  } catch (const std::bad_alloc&) {
    g_last_error = "out of host memory";
    return INS_ERR_NOMEM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return INS_ERR_INTERNAL;
  }
}

insp::Layout to_layout(int layout) {
  if (layout < 0 || layout > 2) throw insp::InvalidArgument("unknown layout");
  return static_cast<insp::Layout>(layout);
}

insp::PoissonSolver& S(ins_poisson_t h) {
  if (!h || !h->s) throw insp::InvalidArgument("null solver handle");
  return *h->s;
}
const insp::HostMesh& M(ins_mesh_t m) {
  if (!m) throw insp::InvalidArgument("null mesh handle");
  return m->m;
}
}  // namespace

extern "C" {

int ins_mesh_generate_square(int nx, int ny, double jitter, uint64_t seed, ins_mesh_t* out) {
  return guarded([&] {
    if (!out || nx < 1 || ny < 1 || jitter < 0 || jitter >= 0.5)
      throw insp::InvalidArgument("invalid square mesh arguments");
    auto* m = new ins_mesh;
    m->m = insp::make_square_mesh(nx, ny, jitter, seed);
    *out = m;
    return INS_OK;
  });
}

int ins_mesh_generate_cube(int n, double jitter, uint64_t seed, ins_mesh_t* out) {
  return guarded([&] {
    if (!out || n < 1 || jitter < 0 || jitter >= 0.5)
      throw insp::InvalidArgument("invalid cube mesh arguments");
    auto* m = new ins_mesh;
    m->m = insp::make_cube_mesh(n, jitter, seed);
    *out = m;
    return INS_OK;
  });
}

int ins_mesh_create(const ins_mesh_desc* d, ins_mesh_t* out) {
  return guarded([&] {
    if (!d || !out || (d->dim != 2 && d->dim != 3) || d->n_nodes < 1 || d->n_cells < 1 ||
        !d->coords || !d->cell2nodes)
      throw insp::InvalidArgument("invalid mesh description");
    const int nv = d->dim + 1;
    for (int64_t i = 0; i < d->n_cells * nv; ++i)
      if (d->cell2nodes[i] < 0 || d->cell2nodes[i] >= d->n_nodes)
        throw insp::InvalidArgument("cell2nodes index out of range");
    auto* m = new ins_mesh;
    m->m.dim = d->dim;
    m->m.n_nodes = d->n_nodes;
    m->m.n_cells = d->n_cells;
    m->m.coords.assign(d->coords, d->coords + d->n_nodes * d->dim);
    m->m.cell2nodes.assign(d->cell2nodes, d->cell2nodes + d->n_cells * nv);
    try {
      insp::build_topology(m->m);
      insp::build_geometry(m->m);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
    return INS_OK;
  });
}

int ins_mesh_info(ins_mesh_t mh, int* dim, int64_t* n_nodes, int64_t* n_cells, int64_t* n_faces,
                  int64_t* n_bfaces) {
  return guarded([&] {
    const auto& m = M(mh);
    if (dim) *dim = m.dim;
    if (n_nodes) *n_nodes = m.n_nodes;
    if (n_cells) *n_cells = m.n_cells;
    if (n_faces) *n_faces = m.n_faces;
    if (n_bfaces) *n_bfaces = m.n_bfaces;
    return INS_OK;
  });
}

int ins_mesh_get(ins_mesh_t mh, double* coords, int* cell2nodes) {
  return guarded([&] {
    const auto& m = M(mh);
    if (coords) std::copy(m.coords.begin(), m.coords.end(), coords);
    if (cell2nodes) std::copy(m.cell2nodes.begin(), m.cell2nodes.end(), cell2nodes);
    return INS_OK;
  });
}

int ins_mesh_get_faces(ins_mesh_t mh, int* face2cells, int* faceNum, int* bface2cells,
                       int* bfaceNum) {
  return guarded([&] {
    const auto& m = M(mh);
    if (face2cells) std::copy(m.face2cells.begin(), m.face2cells.end(), face2cells);
    if (faceNum) std::copy(m.faceNum.begin(), m.faceNum.end(), faceNum);
    if (bface2cells) std::copy(m.bface2cells.begin(), m.bface2cells.end(), bface2cells);
    if (bfaceNum) std::copy(m.bfaceNum.begin(), m.bfaceNum.end(), bfaceNum);
    return INS_OK;
  });
}

void ins_mesh_destroy(ins_mesh_t m) { delete m; }

int ins_poisson_create(ins_mesh_t mesh, int order, ins_poisson_t* out) {
  return guarded([&] {
    if (!out) throw insp::InvalidArgument("null output handle");
    const auto& m = M(mesh);
    auto* h = new ins_poisson;
    try {
      h->s = std::make_unique<insp::PoissonSolver>(m, order);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return INS_OK;
  });
}

void ins_poisson_destroy(ins_poisson_t h) { delete h; }

int ins_poisson_info(ins_poisson_t h, int* np, int* nfp, int64_t* n_cells, int64_t* stride,
                     int64_t* n_bfaces) {
  return guarded([&] {
    auto& s = S(h);
    if (np) *np = s.np();
    if (nfp) *nfp = s.nfp();
    if (n_cells) *n_cells = s.n_cells();
    if (stride) *stride = s.stride();
    if (n_bfaces) *n_bfaces = s.mesh().n_bfaces;
    return INS_OK;
  });
}

int ins_poisson_node_coords(ins_poisson_t h, double* xyz) {
  return guarded([&] {
    if (!xyz) throw insp::InvalidArgument("null output");
    S(h).node_coords(xyz);
    return INS_OK;
  });
}

int ins_poisson_set_bc_types(ins_poisson_t h, const int* bc_types, int64_t n_bfaces) {
  return guarded([&] {
    S(h).set_bc_types(bc_types, n_bfaces);
    return INS_OK;
  });
}

int ins_poisson_set_factor(ins_poisson_t h, const double* factor, int layout) {
  return guarded([&] {
    S(h).set_factor(factor, to_layout(layout));
    return INS_OK;
  });
}

int ins_poisson_set_bcs(ins_poisson_t h, const double* bc_data, int layout) {
  return guarded([&] {
    const auto l = to_layout(layout);
    if (l == insp::Layout::DEVICE_SOA) throw insp::InvalidArgument("bc data must be AoS");
    S(h).set_bcs(bc_data, l);
    return INS_OK;
  });
}

int ins_poisson_set_nullspace(ins_poisson_t h, int nullspace) {
  return guarded([&] {
    S(h).set_nullspace(nullspace != 0);
    return INS_OK;
  });
}

int ins_poisson_set_tol_and_iter(ins_poisson_t h, double rtol, double atol, int max_iter) {
  return guarded([&] {
    S(h).set_tol_and_iter(rtol, atol, max_iter);
    return INS_OK;
  });
}

int ins_poisson_set_solver(ins_poisson_t h, int type) {
  return guarded([&] {
    if (type != INS_SOLVER_JACOBI && type != INS_SOLVER_PMULTIGRID)
      throw insp::InvalidArgument("unsupported solver type");
    S(h).set_solver(static_cast<insp::SolverType>(type));
    return INS_OK;
  });
}

int ins_poisson_set_precision(ins_poisson_t h, int precision) {
  return guarded([&] {
    if (precision != INS_PRECISION_FP64 && precision != INS_PRECISION_MIXED)
      throw insp::InvalidArgument("unsupported precision");
    S(h).set_precision(static_cast<insp::Precision>(precision));
    return INS_OK;
  });
}

int ins_poisson_set_pmg(ins_poisson_t h, const int* orders, int n_levels, const int* pre_it,
                        const int* post_it, const int* cheb_orders, double eig_safety,
                        double cheb_min_factor, int max_iter_eig, int coarse_iters) {
  return guarded([&] {
    if (n_levels < 1 || n_levels > insp::kMaxOrder || !(eig_safety > 0) || !(cheb_min_factor > 0) ||
        !(cheb_min_factor < 1) || max_iter_eig < 1 || coarse_iters < 0)
      throw insp::InvalidArgument("invalid p-multigrid options");
    insp::PMGConfig c;
    for (int i = 0; i < n_levels; ++i) {
      if (orders) c.orders.push_back(orders[i]);
      c.pre_it.push_back(pre_it ? pre_it[i] : 1);
      c.post_it.push_back(post_it ? post_it[i] : 1);
      c.cheb_orders.push_back(cheb_orders ? cheb_orders[i] : 3);
      if (c.pre_it.back() < 0 || c.post_it.back() < 0 || c.cheb_orders.back() < 1)
        throw insp::InvalidArgument("invalid smoother counts");
    }
    c.eig_safety = eig_safety;
    c.cheb_min_factor = cheb_min_factor;
    c.max_iter_eig = max_iter_eig;
    c.coarse_iters = coarse_iters;
    S(h).set_pmg(c);
    return INS_OK;
  });
}

int ins_poisson_init(ins_poisson_t h) {
  return guarded([&] {
    S(h).init();
    return INS_OK;
  });
}

int ins_poisson_solve(ins_poisson_t h, const double* rhs, double* ans, int layout,
                      ins_solve_stats* stats) {
  return guarded([&] {
    if (!rhs || !ans) throw insp::InvalidArgument("null rhs or ans");
    insp::SolveStats st;
    const int rc = S(h).solve(rhs, ans, to_layout(layout), &st);
    if (stats) {
      stats->iterations = st.iterations;
      stats->converged = st.converged;
      stats->rel_residual = st.rel_residual;
      stats->abs_residual = st.abs_residual;
      stats->bnorm = st.bnorm;
      stats->seconds = st.seconds;
      stats->fine_applies = st.fine_applies;
      stats->smoother_applies = st.smoother_applies;
    }
    if (rc == -3) g_last_error = "breakdown: non-positive p.Ap or non-finite residual";
    return rc == 0 ? INS_OK : (rc == 1 ? INS_NOT_CONVERGED : INS_ERR_BREAKDOWN);
  });
}

int ins_poisson_mult(ins_poisson_t h, const double* in, double* out, int layout) {
  return guarded([&] {
    S(h).mult(in, out, to_layout(layout), false);
    return INS_OK;
  });
}

int ins_poisson_mult_fp32(ins_poisson_t h, const double* in, double* out, int layout) {
  return guarded([&] {
    S(h).mult(in, out, to_layout(layout), true);
    return INS_OK;
  });
}

int ins_poisson_diag(ins_poisson_t h, double* diag, int layout) {
  return guarded([&] {
    S(h).diag(diag, to_layout(layout));
    return INS_OK;
  });
}

int ins_poisson_apply_bc(ins_poisson_t h, double* rhs, int layout) {
  return guarded([&] {
    S(h).apply_bc(rhs, to_layout(layout));
    return INS_OK;
  });
}

int ins_poisson_mass(ins_poisson_t h, const double* in, double* out, int layout) {
  return guarded([&] {
    S(h).mass(in, out, to_layout(layout));
    return INS_OK;
  });
}

int ins_poisson_lambda_max(ins_poisson_t h, int level, double* out) {
  return guarded([&] {
    if (!out) throw insp::InvalidArgument("null output");
    *out = S(h).level_lambda_max(level);
    return INS_OK;
  });
}

int ins_poisson_set_profiling(ins_poisson_t h, int on) {
  return guarded([&] {
    S(h).set_profiling(on != 0);
    return INS_OK;
  });
}

int ins_poisson_take_timing(ins_poisson_t h, double* trace_ms, int64_t* trace_launches,
                            double* apply_ms, int64_t* apply_launches) {
  return guarded([&] {
    const insp::KernelTiming t = S(h).take_timing();
    if (trace_ms) *trace_ms = t.trace_ms;
    if (trace_launches) *trace_launches = t.trace_launches;
    if (apply_ms) *apply_ms = t.apply_ms;
    if (apply_launches) *apply_launches = t.apply_launches;
    return INS_OK;
  });
}

int ins_ref_element(int dim, int order, double* rst, double* D, double* M, int* fmask,
                    double* Mface) {
  return guarded([&] {
    if ((dim != 2 && dim != 3) || order < 1 || order > insp::kMaxOrder)
      throw insp::InvalidArgument("unsupported element");
    const insp::RefElement re = insp::make_ref_element(dim, order);
    const int Np = re.Np;
    if (rst) {
      std::copy(re.r.begin(), re.r.end(), rst);
      std::copy(re.s.begin(), re.s.end(), rst + Np);
      if (dim == 3) std::copy(re.t.begin(), re.t.end(), rst + 2 * Np);
    }
    if (D)
      for (int d = 0; d < dim; ++d) std::copy(re.D[d].begin(), re.D[d].end(), D + (size_t)d * Np * Np);
    if (M) std::copy(re.M.begin(), re.M.end(), M);
    if (fmask) std::copy(re.fmask.begin(), re.fmask.end(), fmask);
    if (Mface) std::copy(re.Mface.begin(), re.Mface.end(), Mface);
    return INS_OK;
  });
}

int ins_ref_interp(int dim, int nf, int nc, double* P) {
  return guarded([&] {
    if ((dim != 2 && dim != 3) || nf < 1 || nf > insp::kMaxOrder || nc < 1 ||
        nc > insp::kMaxOrder || !P)
      throw insp::InvalidArgument("invalid interpolation request");
    const insp::RefElement F = insp::make_ref_element(dim, nf), Cc = insp::make_ref_element(dim, nc);
    const std::vector<double> I = insp::interp_matrix(F, Cc);
    std::copy(I.begin(), I.end(), P);
    return INS_OK;
  });
}

const char* ins_poisson_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
