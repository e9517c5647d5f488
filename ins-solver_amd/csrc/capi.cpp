// extern "C" boundary: exceptions never cross it; every entry point returns a
// status code and records a message for ins_poisson_last_error().
#include "../../include/ins_poisson.h"

#include <algorithm>
#include <cmath>
#include <new>
#include <string>

#include <deque>

#include "batch.h"
#include "hdf5_reader.h"
#include "kernels.h"
#include "mesh.h"
#include "oi.h"
#include "refelem.h"
#include "server.h"
#include "viscous_vec.h"
#include "solver.h"

struct ins_mesh {
  insp::HostMesh m;
};
struct ins_poisson {
  std::unique_ptr<insp::PoissonSolver> s;
};
struct ins_viscous_vec {
  std::unique_ptr<insp::VecViscousSolver> s;
};
struct ins_batch {
  std::unique_ptr<insp::BatchEngine> e;
  insp::BatchConfig cfg;
  std::deque<int> finished;
};
struct ins_server {
  std::unique_ptr<insp::Server> s;
  insp::ServerConfig cfg;
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
  } catch (const std::invalid_argument& e) {
    g_last_error = e.what();
    return INS_ERR_INVALID;
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
insp::VecViscousSolver& V(ins_viscous_vec_t h) {
  if (!h || !h->s) throw insp::InvalidArgument("null solver handle");
  return *h->s;
}
void fill_stats(const insp::SolveStats& st, ins_solve_stats* stats) {
  if (!stats) return;
  stats->iterations = st.iterations;
  stats->converged = st.converged;
  stats->rel_residual = st.rel_residual;
  stats->abs_residual = st.abs_residual;
  stats->bnorm = st.bnorm;
  stats->seconds = st.seconds;
  stats->fine_applies = st.fine_applies;
  stats->smoother_applies = st.smoother_applies;
}
const insp::HostMesh& M(ins_mesh_t m) {
  if (!m) throw insp::InvalidArgument("null mesh handle");
  return m->m;
}

insp::BatchConfig to_batch_config(int dim, const ins_batch_config* c) {
  insp::BatchConfig b;
  b.dim = dim;
  if (!c) return b;
  if (c->precision != INS_PRECISION_FP64 && c->precision != INS_PRECISION_MIXED)
    throw insp::InvalidArgument("unknown precision");
  if (c->cheb_order < 1 || c->cheb_order > 8 || c->pre_it < 1 || c->post_it < 0 ||
      c->eig_safety <= 0 || c->cheb_min_factor <= 0 || c->cheb_min_factor >= 1 ||
      c->max_iter_eig < 1 || c->coarse_iters < 0)
    throw insp::InvalidArgument("invalid batch configuration");
  b.fp32 = c->precision == INS_PRECISION_MIXED;
  b.cheb_order = c->cheb_order;
  b.pre_it = c->pre_it;
  b.post_it = c->post_it;
  b.eig_safety = c->eig_safety;
  b.cheb_min_factor = c->cheb_min_factor;
  b.max_iter_eig = c->max_iter_eig;
  b.coarse_iters = c->coarse_iters;
  b.use_graph = c->use_graph != 0;
  if (c->coarse != INS_COARSE_JACOBI_CG && c->coarse != INS_COARSE_AMG)
    throw insp::InvalidArgument("unknown coarse solve");
  b.amg = c->coarse == INS_COARSE_AMG;
  if (c->tile != 0 && c->tile != 64 && c->tile != 128 && c->tile != 256)
    throw insp::InvalidArgument("tile must be 64, 128, 256 or 0");
  b.tile = c->tile;
  if (c->chunk < 0 || c->chunk > 64) throw insp::InvalidArgument("chunk must be in 1..64 (0 = default)");
  if (c->chunk > 0) b.chunk = c->chunk;
  return b;
}

int np_of(int dim, int N) { return dim == 2 ? (N + 1) * (N + 2) / 2 : (N + 1) * (N + 2) * (N + 3) / 6; }

// Copies one request into a problem spec. `own` copies the mesh (the request
// outlives the call); otherwise the spec borrows it for the call's duration.
insp::ProblemSpec to_spec(const ins_problem& p, bool own) {
  if (!p.mesh) throw insp::InvalidArgument("problem without a mesh");
  if (p.order < 1 || p.order > insp::kMaxOrder) throw insp::InvalidArgument("order must be in 1..4");
  if (!p.rhs) throw insp::InvalidArgument("problem without a rhs");
  if (!(p.rtol >= 0) || !(p.atol >= 0) || p.max_iter < 1)
    throw insp::InvalidArgument("invalid tolerance or iteration limit");
  const insp::HostMesh& m = p.mesh->m;
  insp::ProblemSpec sp;
  sp.mesh = own ? std::make_shared<insp::HostMesh>(m)
                : insp::HostMeshPtr(insp::HostMeshPtr(), const_cast<insp::HostMesh*>(&m));
  sp.order = p.order;
  const int64_t n = m.n_cells * np_of(m.dim, p.order);
  if (p.bc_types) {
    for (int64_t i = 0; i < m.n_bfaces; ++i)
      if (p.bc_types[i] != INS_BC_DIRICHLET && p.bc_types[i] != INS_BC_NEUMANN)
        throw insp::InvalidArgument("unknown boundary type");
    sp.bc_types.assign(p.bc_types, p.bc_types + m.n_bfaces);
  }
  if (p.factor) sp.factor.assign(p.factor, p.factor + n);
  sp.nullspace = p.nullspace != 0;
  sp.rtol = p.rtol;
  sp.atol = p.atol;
  sp.max_iter = p.max_iter;
  sp.rhs.assign(p.rhs, p.rhs + n);
  sp.rhs_nodal = p.rhs_nodal != 0;
  return sp;
}

int fill_result(const insp::BatchConfig& cfg, const insp::ProbResult& r, double* x,
                ins_solve_stats* st) {
  if (r.status == insp::ProbStatus::RUNNING) throw insp::InvalidArgument("problem is still running");
  if (x) std::copy(r.x.begin(), r.x.end(), x);
  if (st) {
    st->iterations = r.iterations;
    st->converged = r.status == insp::ProbStatus::CONVERGED;
    st->rel_residual = r.rel_residual;
    st->abs_residual = r.rel_residual * r.bnorm;
    st->bnorm = r.bnorm;
    st->seconds = r.seconds;
    st->fine_applies = r.fine_applies;
    st->smoother_applies =
        r.vcycles * ((int64_t)(cfg.pre_it + cfg.post_it) * cfg.cheb_order + cfg.coarse_iters);
  }
  switch (r.status) {
    case insp::ProbStatus::CONVERGED: return INS_OK;
    case insp::ProbStatus::NOT_CONVERGED:
      g_last_error = "not converged";
      return INS_NOT_CONVERGED;
    default:
      g_last_error = "breakdown: p.Ap <= 0 or a non-finite residual";
      return INS_ERR_BREAKDOWN;
  }
}

insp::BatchEngine& B(ins_batch_t b) {
  if (!b || !b->e) throw insp::InvalidArgument("null batch handle");
  return *b->e;
}
insp::Server& SV(ins_server_t s) {
  if (!s || !s->s) throw insp::InvalidArgument("null server handle");
  return *s->s;
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

namespace {
// A mesh with the caller's OP2 face sets (ins_mesh_create_op2, ins_mesh_load_op2_hdf5).
ins_mesh* make_op2_mesh(int dim, int64_t n_nodes, const double* coords, int64_t n_cells,
                        const int* cell2nodes, int64_t n_faces, const int* face2cells,
                        const int* faceNum, int64_t n_bfaces, const int* bface2cells,
                        const int* bfaceNum) {
  if ((dim != 2 && dim != 3) || n_nodes < 1 || n_cells < 1 || !coords || !cell2nodes || n_faces < 0 ||
      n_bfaces < 0)
    throw insp::InvalidArgument("invalid mesh description");
  const int nv = dim + 1;
  for (int64_t i = 0; i < n_cells * nv; ++i)
    if (cell2nodes[i] < 0 || cell2nodes[i] >= n_nodes)
      throw insp::InvalidArgument("cell2nodes index out of range");
  auto m = std::make_unique<ins_mesh>();
  m->m.dim = dim;
  m->m.n_nodes = n_nodes;
  m->m.n_cells = n_cells;
  m->m.coords.assign(coords, coords + n_nodes * dim);
  m->m.cell2nodes.assign(cell2nodes, cell2nodes + n_cells * nv);
  // The caller's local face numbers and nodal data refer to its own vertex order, so a
  // negatively oriented cell cannot be re-oriented here: it is rejected.
  const int64_t neg = insp::first_negative_cell(m->m);
  if (neg >= 0)
    throw insp::InvalidArgument("cell " + std::to_string(neg) +
                                " is negatively oriented (its vertices must be listed counter-"
                                "clockwise in 2D, with positive volume in 3D)");
  insp::build_topology(m->m);
  insp::adopt_op2_faces(m->m, n_faces, face2cells, faceNum, n_bfaces, bface2cells, bfaceNum);
  insp::build_geometry(m->m);
  return m.release();
}
}  // namespace

int ins_mesh_create_op2(const ins_mesh_desc* d, const ins_mesh_faces_desc* f, ins_mesh_t* out) {
  return guarded([&] {
    if (!d || !f || !out) throw insp::InvalidArgument("invalid mesh description");
    *out = make_op2_mesh(d->dim, d->n_nodes, d->coords, d->n_cells, d->cell2nodes, f->n_faces,
                         f->face2cells, f->faceNum, f->n_bfaces, f->bface2cells, f->bfaceNum);
    return INS_OK;
  });
}

int ins_mesh_load_op2_hdf5(const char* path, ins_mesh_t* out) {
  return guarded([&] {
    if (!path || !out) throw insp::InvalidArgument("null path or output");
    const insp::Hdf5File f(path);
    auto pick = [&](const char* a, const char* b) -> std::string {
      if (f.has(a)) return a;
      if (f.has(b)) return b;
      return "";
    };
    auto set_size = [&](const char* a, const char* b) -> int64_t {
      const std::string n = pick(a, b);
      if (n.empty()) return -1;
      const std::vector<int32_t> v = f.read_int(n);
      if (v.size() != 1) throw insp::InvalidArgument(std::string(path) + ": set " + n + " is not a size");
      return v[0];
    };
    std::vector<uint64_t> dc, dm;
    const std::vector<double> coords = f.read_double("node_coords", &dc);
    const std::vector<int32_t> c2n = f.read_int("cell2nodes", &dm);
    if (dc.size() != 2 || (dc[1] != 2 && dc[1] != 3))
      throw insp::InvalidArgument(std::string(path) + ": node_coords must be nodes x 2 or nodes x 3");
    const int dim = (int)dc[1];
    if (dm.size() != 2 || dm[1] != (uint64_t)dim + 1)
      throw insp::InvalidArgument(std::string(path) + ": cell2nodes must be cells x (dim+1)");
    const int64_t n_nodes = (int64_t)dc[0], n_cells = (int64_t)dm[0];
    if (set_size("nodes", "nodes") >= 0 && set_size("nodes", "nodes") != n_nodes)
      throw insp::InvalidArgument(std::string(path) + ": set nodes disagrees with node_coords");
    if (set_size("cells", "cells") >= 0 && set_size("cells", "cells") != n_cells)
      throw insp::InvalidArgument(std::string(path) + ": set cells disagrees with cell2nodes");
    const std::string f2c_name = pick("face2cells", "edge2cells"), fnum_name = pick("faceNum", "edgeNum");
    if (f2c_name.empty() || fnum_name.empty())
      throw insp::InvalidArgument(std::string(path) + ": no face2cells / faceNum (edgeNum) maps");
    std::vector<uint64_t> df, dn;
    const std::vector<int32_t> f2c = f.read_int(f2c_name, &df), fnum = f.read_int(fnum_name, &dn);
    if (df.size() != 2 || df[1] != 2 || dn != df)
      throw insp::InvalidArgument(std::string(path) + ": face maps must be faces x 2");
    const int64_t n_faces = (int64_t)df[0];
    const std::string b2c_name = pick("bface2cells", "bedge2cells"), bnum_name = pick("bfaceNum", "bedgeNum");
    std::vector<int32_t> b2c, bnum;
    if (!b2c_name.empty() && !bnum_name.empty()) {
      b2c = f.read_int(b2c_name);
      bnum = f.read_int(bnum_name);
      if (b2c.size() != bnum.size())
        throw insp::InvalidArgument(std::string(path) + ": bface maps differ in size");
    } else if (set_size("bfaces", "bedges") > 0) {
      throw insp::InvalidArgument(std::string(path) + ": boundary faces without bface2cells / bfaceNum");
    }
    *out = make_op2_mesh(dim, n_nodes, coords.data(), n_cells, c2n.data(), n_faces, f2c.data(),
                         fnum.data(), (int64_t)b2c.size(), b2c.data(), bnum.data());
    return INS_OK;
  });
}

int ins_mesh_renumber(ins_mesh_t mh, int method, int64_t* perm) {
  return guarded([&] {
    if (!mh) throw insp::InvalidArgument("null mesh");
    if (method != INS_RENUMBER_SHAPE_CLASS && method != INS_RENUMBER_SHAPE_CLASS_WINDOWED &&
        method != INS_RENUMBER_SHAPE_CLASS_TILED)
      throw insp::InvalidArgument("unsupported renumbering");
    const std::vector<int64_t> p = insp::shape_class_order(
        mh->m, method == INS_RENUMBER_SHAPE_CLASS ? 0 : (method == INS_RENUMBER_SHAPE_CLASS_WINDOWED ? 1 : 2));
    mh->m = insp::renumber_cells(mh->m, p);
    if (perm) std::copy(p.begin(), p.end(), perm);
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

int ins_poisson_set_factor_oi(ins_poisson_t h, const double* factor_vol, int layout) {
  return guarded([&] {
    S(h).set_factor_oi(factor_vol, to_layout(layout));
    return INS_OK;
  });
}

int ins_poisson_set_factor_surf_oi(ins_poisson_t h, const double* factor_surf, int layout) {
  return guarded([&] {
    S(h).set_factor_surf_oi(factor_surf, to_layout(layout));
    return INS_OK;
  });
}

int ins_oi_rule(int dim, int order, int* n_vol, int* n_surf, double* vol_rst, double* vol_w,
                double* surf_bary, double* surf_w) {
  return guarded([&] {
    if (dim != 2 && dim != 3) throw insp::InvalidArgument("dimension must be 2 or 3");
    if (order < 1 || order > insp::kMaxOrder) throw insp::InvalidArgument("order must be in 1..4");
    const insp::OiRule q = insp::make_oi_rule(dim, order);
    if (n_vol) *n_vol = q.nv;
    if (n_surf) *n_surf = q.ns;
    if (vol_rst) std::copy(q.vx.begin(), q.vx.end(), vol_rst);
    if (vol_w) std::copy(q.vw.begin(), q.vw.end(), vol_w);
    if (surf_bary) std::copy(q.sb.begin(), q.sb.end(), surf_bary);
    if (surf_w) std::copy(q.sw.begin(), q.sw.end(), surf_w);
    return INS_OK;
  });
}

int ins_oi_interp(int dim, int order, double* vol, double* surf) {
  return guarded([&] {
    if (dim != 2 && dim != 3) throw insp::InvalidArgument("dimension must be 2 or 3");
    if (order < 1 || order > insp::kMaxOrder) throw insp::InvalidArgument("order must be in 1..4");
    const insp::OiRule q = insp::make_oi_rule(dim, order);
    const insp::RefElement re = insp::make_ref_element(dim, order);
    std::vector<double> Iv, Is, xv, xs;
    insp::oi_interp(re, q, Iv, Is, xv, xs);
    if (vol) std::copy(Iv.begin(), Iv.end(), vol);
    if (surf) std::copy(Is.begin(), Is.end(), surf);
    return INS_OK;
  });
}

int ins_poisson_set_mm_factor(ins_poisson_t h, const double* mm, int layout) {
  return guarded([&] {
    S(h).set_mm_factor(mm, to_layout(layout));
    return INS_OK;
  });
}

int ins_poisson_set_mm_const(ins_poisson_t h, double mm) {
  return guarded([&] {
    S(h).set_mm_const(mm);
    return INS_OK;
  });
}

int ins_poisson_set_inv_mass_factor(ins_poisson_t h, double f) {
  return guarded([&] {
    if (!std::isfinite(f) || f == 0.0) throw insp::InvalidArgument("inverse-mass factor must be finite and nonzero");
    S(h).set_inv_mass_factor(f);
    return INS_OK;
  });
}

int ins_poisson_set_inv_mass_recp_factor(ins_poisson_t h, const double* f, int layout) {
  return guarded([&] {
    S(h).set_inv_mass_recp_factor(f, to_layout(layout));
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
    if (type != INS_SOLVER_JACOBI && type != INS_SOLVER_BLOCK_JACOBI && type != INS_SOLVER_INV_MASS &&
        type != INS_SOLVER_PMULTIGRID)
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

int ins_poisson_set_element_pass(ins_poisson_t h, int kind) {
  return guarded([&] {
    if (kind != INS_ELEMENT_PASS_AUTO && kind != INS_ELEMENT_PASS_VECTOR && kind != INS_ELEMENT_PASS_MATRIX)
      throw insp::InvalidArgument("unsupported element pass");
    S(h).set_element_pass(static_cast<insp::ElementPass>(kind));
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
    c.coarse = S(h).pmg().coarse;  // the coarse solve kind is set by ins_poisson_set_coarse
    c.amg_cycles = S(h).pmg().amg_cycles;
    S(h).set_pmg(c);
    return INS_OK;
  });
}

int ins_poisson_inv_block_diag(ins_poisson_t h, double* out, int layout) {
  return guarded([&] {
    if (!out || layout != INS_LAYOUT_HOST_AOS) throw insp::InvalidArgument("host output required");
    S(h).inv_block_diag(out);
    return INS_OK;
  });
}

int ins_poisson_pressure_rhs(ins_poisson_t h, double b0, double b1, double dt, const double* dPdN,
                             double* dPdN_old, double* div, int layout) {
  return guarded([&] {
    S(h).pressure_rhs(b0, b1, dt, dPdN, dPdN_old, div, to_layout(layout));
    return INS_OK;
  });
}

int ins_poisson_set_history_size(ins_poisson_t h, int n) {
  return guarded([&] {
    S(h).set_history_size(n);
    return INS_OK;
  });
}

int ins_poisson_history_push(ins_poisson_t h, double t, const double* x, int layout) {
  return guarded([&] {
    S(h).history_push(t, x, to_layout(layout));
    return INS_OK;
  });
}

int ins_poisson_initial_guess(ins_poisson_t h, double t, double* ans, int layout) {
  return guarded([&] {
    S(h).extrapolate(t, ans, to_layout(layout));
    return INS_OK;
  });
}

int ins_poisson_set_coarse(ins_poisson_t h, int kind, int amg_cycles) {
  return guarded([&] {
    auto& s = S(h);
    if ((kind != INS_COARSE_JACOBI_CG && kind != INS_COARSE_AMG) || amg_cycles < 1 || amg_cycles > 16)
      throw insp::InvalidArgument("invalid coarse solve settings");
    insp::PMGConfig c = s.pmg();
    c.coarse = kind == INS_COARSE_AMG ? insp::CoarseType::AMG : insp::CoarseType::JACOBI_CG;
    c.amg_cycles = amg_cycles;
    s.set_pmg(c);
    return INS_OK;
  });
}

int ins_poisson_coarse_matrix(ins_poisson_t h, int64_t* n, int64_t* nnz, int* rowptr, int* cols,
                              double* vals) {
  return guarded([&] {
    const insp::HostCsr& A = S(h).vertex_matrix();
    if (n) *n = A.n;
    if (nnz) *nnz = A.nnz();
    if (rowptr) std::copy(A.ptr.begin(), A.ptr.end(), rowptr);
    if (cols) std::copy(A.col.begin(), A.col.end(), cols);
    if (vals) std::copy(A.val.begin(), A.val.end(), vals);
    return INS_OK;
  });
}

int ins_poisson_amg_levels(ins_poisson_t h, int* n_levels, int* sizes) {
  return guarded([&] {
    const std::vector<int> z = S(h).amg_level_sizes();
    if (!n_levels) throw insp::InvalidArgument("null output");
    *n_levels = (int)z.size();
    if (sizes)
      for (size_t i = 0; i < z.size() && i < 16; ++i) sizes[i] = z[i];
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

int ins_poisson_set_dat_stride(ins_poisson_t h, int64_t stride) {
  return guarded([&] {
    S(h).set_dat_stride(stride);
    return INS_OK;
  });
}

int ins_poisson_mult_noalloc(ins_poisson_t h, const double* in, double* out, int layout) {
  static_assert(INS_NOALLOC_SCRATCH_BYTES == (long long)insp::kNoAllocScratchBytes,
                "scratch size of the header and the kernels");
  return guarded([&] {
    S(h).mult_noalloc(in, out, to_layout(layout));
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

int ins_poisson_take_coarse_timing(ins_poisson_t h, double* apply_ms, int64_t* launches) {
  return guarded([&] {
    const insp::CoarseTiming t = S(h).take_coarse_timing();
    if (apply_ms) *apply_ms = t.apply_ms;
    if (launches) *launches = t.launches;
    return INS_OK;
  });
}

int ins_poisson_grad_central(ins_poisson_t h, const double* p, double* dpdx, double* dpdy, double* dpdz,
                             int layout) {
  return guarded([&] {
    double* const g[3] = {dpdx, dpdy, dpdz};
    S(h).grad_central(p, g, nullptr, 0.0, to_layout(layout));
    return INS_OK;
  });
}

int ins_poisson_project_velocity(ins_poisson_t h, double dt, const double* p, const double* uTx,
                                 const double* uTy, const double* uTz, double* uTTx, double* uTTy,
                                 double* uTTz, int layout) {
  return guarded([&] {
    double* const g[3] = {uTTx, uTTy, uTTz};
    const double* const u[3] = {uTx, uTy, uTz};
    S(h).grad_central(p, g, u, dt, to_layout(layout));
    return INS_OK;
  });
}

int ins_device_peaks(double* copy_gbs, double* mfma_f32_tflops, double* mfma_f64_tflops) {
  return guarded([&] {
    insp::measure_device_peaks(copy_gbs, mfma_f32_tflops, mfma_f64_tflops);
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

int ins_mesh_node_coords(ins_mesh_t mh, int order, double* xyz) {
  return guarded([&] {
    const auto& m = M(mh);
    if (order < 1 || order > insp::kMaxOrder || !xyz) throw insp::InvalidArgument("invalid order or output");
    insp::mesh_node_coords(m, insp::make_ref_element(m.dim, order), xyz);
    return INS_OK;
  });
}

void ins_batch_config_default(ins_batch_config* c) {
  if (!c) return;
  const insp::BatchConfig d;
  c->precision = d.fp32 ? INS_PRECISION_MIXED : INS_PRECISION_FP64;
  c->cheb_order = d.cheb_order;
  c->pre_it = d.pre_it;
  c->post_it = d.post_it;
  c->eig_safety = d.eig_safety;
  c->cheb_min_factor = d.cheb_min_factor;
  c->max_iter_eig = d.max_iter_eig;
  c->coarse_iters = d.coarse_iters;
  c->use_graph = d.use_graph ? 1 : 0;
  c->coarse = d.amg ? INS_COARSE_AMG : INS_COARSE_JACOBI_CG;
  c->tile = d.tile;
  c->chunk = d.chunk;
}

int ins_batch_setup_seconds(ins_batch_t b, double* t5) {
  return guarded([&] {
    auto& e = B(b);
    if (!t5) throw insp::InvalidArgument("null output");
    std::copy(e.setup_seconds(), e.setup_seconds() + 5, t5);
    return INS_OK;
  });
}

int ins_batch_tuning(ins_batch_t b, int* tile, double* ms) {
  return guarded([&] {
    auto& e = B(b);
    if (tile) *tile = e.tile();
    if (ms) {
      const insp::BatchEngine::TuneResult r = insp::BatchEngine::last_tune(b->cfg.dim, b->cfg.fp32);
      const bool tuned = b->cfg.tile == 0 && r.tile != 0;
      for (int i = 0; i < 3; ++i) ms[i] = tuned ? r.ms[i] : 0.0;
    }
    return INS_OK;
  });
}

int ins_batch_amg_bytes(int dim, int64_t n_nodes, int64_t n_cells, int precision, int64_t* bytes) {
  return guarded([&] {
    if ((dim != 2 && dim != 3) || n_nodes < 1 || n_cells < 1 || !bytes)
      throw insp::InvalidArgument("invalid problem size");
    *bytes = (int64_t)insp::BatchEngine::amg_bytes_estimate(dim, n_nodes, n_cells,
                                                            precision == INS_PRECISION_MIXED);
    return INS_OK;
  });
}

int ins_batch_problem_bytes(int dim, int64_t n_cells, int order, int has_factor, int precision,
                            int64_t* bytes) {
  return guarded([&] {
    if ((dim != 2 && dim != 3) || n_cells < 1 || order < 1 || order > insp::kMaxOrder || !bytes)
      throw insp::InvalidArgument("invalid problem size");
    *bytes = (int64_t)insp::BatchEngine::arena_bytes(dim, n_cells, order, has_factor != 0,
                                                     precision == INS_PRECISION_MIXED);
    return INS_OK;
  });
}

int ins_batch_create(int dim, const ins_batch_config* cfg, ins_batch_t* out) {
  return guarded([&] {
    if (!out || (dim != 2 && dim != 3)) throw insp::InvalidArgument("invalid batch arguments");
    auto h = std::make_unique<ins_batch>();
    h->cfg = to_batch_config(dim, cfg);
    h->e = std::make_unique<insp::BatchEngine>(h->cfg);
    *out = h.release();
    return INS_OK;
  });
}

void ins_batch_destroy(ins_batch_t b) { delete b; }

int ins_batch_add(ins_batch_t b, const ins_problem* probs, int n, int* ids) {
  return guarded([&] {
    auto& e = B(b);
    if (n < 0 || (n > 0 && (!probs || !ids))) throw insp::InvalidArgument("invalid problem list");
    std::vector<insp::ProblemSpec> specs;
    for (int i = 0; i < n; ++i) {
      if (!probs[i].mesh || probs[i].mesh->m.dim != b->cfg.dim)
        throw insp::InvalidArgument("problem dimension does not match the batch");
      specs.push_back(to_spec(probs[i], false));
    }
    const std::vector<int> got = e.add_many(std::move(specs));
    std::copy(got.begin(), got.end(), ids);
    return INS_OK;
  });
}

int ins_batch_iterate(ins_batch_t b, int max_iters, int* running) {
  return guarded([&] {
    auto& e = B(b);
    if (max_iters < 0) throw insp::InvalidArgument("negative iteration count");
    // chunks of lockstep iterations, one host round trip each
    for (int done = 0; done < max_iters && e.running() > 0;)
      done += e.iterate(std::min(std::max(1, e.config().chunk), max_iters - done));
    if (running) *running = e.running();
    return INS_OK;
  });
}

int ins_batch_take_finished(ins_batch_t b, int* ids, int cap, int* n) {
  return guarded([&] {
    auto& e = B(b);
    if (cap < 0 || (cap > 0 && !ids) || !n) throw insp::InvalidArgument("invalid output buffer");
    for (int id : e.take_finished()) b->finished.push_back(id);
    int k = 0;
    while (k < cap && !b->finished.empty()) {
      ids[k++] = b->finished.front();
      b->finished.pop_front();
    }
    *n = k;
    return INS_OK;
  });
}

int ins_batch_result(ins_batch_t b, int id, double* x, ins_solve_stats* stats) {
  return guarded([&] { return fill_result(b ? b->cfg : insp::BatchConfig{}, B(b).result(id), x, stats); });
}

int ins_batch_release(ins_batch_t b, int id) {
  return guarded([&] {
    B(b).release(id);
    return INS_OK;
  });
}

int ins_batch_stats(ins_batch_t b, int64_t* bytes_in_use, int64_t* iterations_run, int* running) {
  return guarded([&] {
    auto& e = B(b);
    if (bytes_in_use) *bytes_in_use = (int64_t)e.bytes_in_use();
    if (iterations_run) *iterations_run = e.iterations_run();
    if (running) *running = e.running();
    return INS_OK;
  });
}

int ins_server_create(int64_t hbm_budget_bytes, const ins_batch_config* cfg2d,
                      const ins_batch_config* cfg3d, ins_server_t* out) {
  return guarded([&] {
    if (!out || hbm_budget_bytes <= 0) throw insp::InvalidArgument("invalid server arguments");
    auto h = std::make_unique<ins_server>();
    h->cfg.hbm_budget = (size_t)hbm_budget_bytes;
    h->cfg.cfg2d = to_batch_config(2, cfg2d);
    h->cfg.cfg3d = to_batch_config(3, cfg3d);
    h->s = std::make_unique<insp::Server>(h->cfg);
    *out = h.release();
    return INS_OK;
  });
}

void ins_server_destroy(ins_server_t s) { delete s; }

int ins_server_submit(ins_server_t s, const ins_problem* p, int64_t* id) {
  return guarded([&] {
    auto& sv = SV(s);
    if (!p || !id) throw insp::InvalidArgument("invalid request");
    *id = sv.submit(to_spec(*p, true));
    return INS_OK;
  });
}

int ins_server_poll(ins_server_t s, int64_t id, int* done) {
  return guarded([&] {
    auto& sv = SV(s);
    if (!done) throw insp::InvalidArgument("null output");
    *done = sv.poll(id) ? 1 : 0;
    return INS_OK;
  });
}

int ins_server_wait(ins_server_t s, int64_t id, double* x, ins_solve_stats* stats, double* latency_s,
                    double* queue_s) {
  return guarded([&] {
    auto& sv = SV(s);
    insp::RequestResult r = sv.wait(id);
    if (latency_s) *latency_s = r.latency_s;
    if (queue_s) *queue_s = r.queue_s;
    if (r.rejected) {
      g_last_error = "request needs more device memory than the server budget";
      return (int)INS_ERR_NOMEM;
    }
    const insp::ServerConfig& c = s->cfg;
    return fill_result(r.dim == 3 ? c.cfg3d : c.cfg2d, r.result, x, stats);
  });
}

int ins_server_wait_all(ins_server_t s) {
  return guarded([&] {
    SV(s).wait_all();
    return INS_OK;
  });
}

int ins_server_stats(ins_server_t s, int64_t* completed, int64_t* peak_bytes) {
  return guarded([&] {
    auto& sv = SV(s);
    if (completed) *completed = sv.completed();
    if (peak_bytes) *peak_bytes = (int64_t)sv.peak_bytes();
    return INS_OK;
  });
}

int ins_viscous_vec_create(ins_mesh_t mesh, int order, ins_viscous_vec_t* out) {
  return guarded([&] {
    if (!out) throw insp::InvalidArgument("null output handle");
    const auto& m = M(mesh);
    auto* h = new ins_viscous_vec;
    try {
      h->s = std::make_unique<insp::VecViscousSolver>(m, order);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return INS_OK;
  });
}

void ins_viscous_vec_destroy(ins_viscous_vec_t h) { delete h; }

int ins_viscous_vec_info(ins_viscous_vec_t h, int* ncomp, int* np, int* nfp, int64_t* n_cells,
                         int64_t* stride, int64_t* n_bfaces) {
  return guarded([&] {
    auto& v = V(h);
    auto& s = v.base();
    if (ncomp) *ncomp = v.ncomp();
    if (np) *np = s.np();
    if (nfp) *nfp = s.nfp();
    if (n_cells) *n_cells = s.n_cells();
    if (stride) *stride = s.stride();
    if (n_bfaces) *n_bfaces = s.mesh().n_bfaces;
    return INS_OK;
  });
}

int ins_viscous_vec_set_dat_stride(ins_viscous_vec_t h, int64_t stride) {
  return guarded([&] {
    V(h).base().set_dat_stride(stride);
    return INS_OK;
  });
}

int ins_viscous_vec_set_bc_types(ins_viscous_vec_t h, const int* const* bc_types, int64_t n_bfaces) {
  return guarded([&] {
    V(h).set_bc_types(bc_types, n_bfaces);
    return INS_OK;
  });
}

int ins_viscous_vec_set_factor(ins_viscous_vec_t h, const double* factor, int layout) {
  return guarded([&] {
    auto& v = V(h);
    v.base().set_factor(factor, to_layout(layout));
    v.touch();
    return INS_OK;
  });
}

int ins_viscous_vec_set_mm_factor(ins_viscous_vec_t h, const double* mm, int layout) {
  return guarded([&] {
    auto& v = V(h);
    v.base().set_mm_factor(mm, to_layout(layout));
    v.touch();
    return INS_OK;
  });
}

int ins_viscous_vec_set_mm_const(ins_viscous_vec_t h, double mm) {
  return guarded([&] {
    auto& v = V(h);
    v.base().set_mm_const(mm);
    v.touch();
    return INS_OK;
  });
}

int ins_viscous_vec_set_bcs(ins_viscous_vec_t h, const double* const* bc_data, int layout) {
  return guarded([&] {
    const auto l = to_layout(layout);
    if (l == insp::Layout::DEVICE_SOA) throw insp::InvalidArgument("bc data must be AoS");
    V(h).set_bcs(bc_data, l);
    return INS_OK;
  });
}

int ins_viscous_vec_set_tol_and_iter(ins_viscous_vec_t h, double rtol, double atol, int max_iter) {
  return guarded([&] {
    V(h).set_tol_and_iter(rtol, atol, max_iter);
    return INS_OK;
  });
}

int ins_viscous_vec_set_preconditioner(ins_viscous_vec_t h, int preconditioner) {
  return guarded([&] {
    if (preconditioner < INS_VEC_PRE_NONE || preconditioner > INS_VEC_PRE_BLOCK_JACOBI)
      throw insp::InvalidArgument("unknown preconditioner");
    V(h).set_preconditioner(static_cast<insp::VecPrecond>(preconditioner));
    return INS_OK;
  });
}

int ins_viscous_vec_set_inv_mass_factor(ins_viscous_vec_t h, double f) {
  return guarded([&] {
    if (!std::isfinite(f)) throw insp::InvalidArgument("inverse-mass factor must be finite");
    V(h).base().set_inv_mass_factor(f);
    return INS_OK;
  });
}

int ins_viscous_vec_set_inv_mass_recp_factor(ins_viscous_vec_t h, const double* f, int layout) {
  return guarded([&] {
    V(h).base().set_inv_mass_recp_factor(f, to_layout(layout));
    return INS_OK;
  });
}

int ins_viscous_vec_mult(ins_viscous_vec_t h, const double* const* in, double* const* out, int layout) {
  return guarded([&] {
    if (!in || !out) throw insp::InvalidArgument("null vector");
    V(h).mult(in, out, to_layout(layout));
    return INS_OK;
  });
}

int ins_viscous_vec_apply_bc(ins_viscous_vec_t h, double* const* rhs, int layout) {
  return guarded([&] {
    if (!rhs) throw insp::InvalidArgument("null vector");
    V(h).apply_bc(rhs, to_layout(layout));
    return INS_OK;
  });
}

int ins_viscous_vec_diag(ins_viscous_vec_t h, double* const* diag, int layout) {
  return guarded([&] {
    if (!diag) throw insp::InvalidArgument("null vector");
    V(h).diag(diag, to_layout(layout));
    return INS_OK;
  });
}

int ins_viscous_vec_inv_block_diag(ins_viscous_vec_t h, double* out) {
  return guarded([&] {
    if (!out) throw insp::InvalidArgument("null output");
    V(h).inv_block_diag(out);
    return INS_OK;
  });
}

int ins_viscous_vec_solve(ins_viscous_vec_t h, const double* const* rhs, double* const* ans, int layout,
                          ins_solve_stats* stats) {
  return guarded([&] {
    if (!rhs || !ans) throw insp::InvalidArgument("null rhs or ans");
    insp::SolveStats st;
    const int rc = V(h).solve(rhs, ans, to_layout(layout), &st);
    fill_stats(st, stats);
    return rc;
  });
}

const char* ins_poisson_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
