// Unstructured simplex meshes for the DG Poisson solve: generation of the
// jittered benchmark meshes, face connectivity (the OP2 sets faces/bfaces and
// maps face2cells/bface2cells with local face numbers), and per-element
// geometric metrics of straight-sided simplices.
#pragma once

#include <cstdint>
#include <vector>

namespace insp {

struct HostMesh {
  int dim = 0;
  int64_t n_nodes = 0, n_cells = 0;
  std::vector<double> coords;   // n_nodes x dim
  std::vector<int> cell2nodes;  // n_cells x (dim+1), positively oriented

  // Derived connectivity (build_topology).
  std::vector<int> nbr;         // n_cells x Nf: neighbour cell, or -(bface+1)
  std::vector<int> nbr_face;    // n_cells x Nf: neighbour's local face (0 on boundary)
  std::vector<int> nbr_perm;    // n_cells x Nf: vertex permutation index (0 on boundary)
  int64_t n_faces = 0;          // interior faces
  std::vector<int> face2cells;  // n_faces x 2 (left cell < right cell)
  std::vector<int> faceNum;     // n_faces x 2 local face numbers
  int64_t n_bfaces = 0;
  std::vector<int> bface2cells; // n_bfaces
  std::vector<int> bfaceNum;    // n_bfaces
  int64_t n_periodic = 0;       // interior faces joining two boundary faces by a translation

  // Geometry (build_geometry): per cell J and the metric K[r][s] = grad(xi_r).grad(xi_s)
  // scaled by J (Ktilde), stored as 3 (2D: rr, rs, ss) or 6 (3D: rr, rs, rt, ss, st, tt).
  std::vector<double> J;        // n_cells
  std::vector<double> Kt;       // n_cells x (3|6)
};

// Unit square, nx*ny cells, each split into two triangles along the (i,j)-(i+1,j+1)
// diagonal; interior vertices jittered uniformly by +-jitter*h per coordinate.
HostMesh make_square_mesh(int nx, int ny, double jitter, uint64_t seed);
// Unit cube, n^3 hexes, each split into 6 Kuhn tetrahedra sharing the main
// diagonal; interior vertices jittered by +-jitter*h per coordinate.
HostMesh make_cube_mesh(int n, double jitter, uint64_t seed);

// Orients cells positively, then builds faces and neighbour tables.
void build_topology(HostMesh& m);
// Index of the first cell with a negative signed measure, or -1.
int64_t first_negative_cell(const HostMesh& m);
// Adopts a caller's OP2 face sets (face2cells/faceNum, bface2cells/bfaceNum, in the caller's
// order) after build_topology: every face must be one of the mesh's faces with the same local
// face numbers (either side first), or a periodic pair of two boundary faces whose vertices
// match up to one translation, and each exactly once. The boundary-face numbering then
// follows the caller's bface order (bc_types and bc_data are indexed by it).
void adopt_op2_faces(HostMesh& m, int64_t n_faces, const int* face2cells, const int* faceNum,
                     int64_t n_bfaces, const int* bface2cells, const int* bfaceNum);
void build_geometry(HostMesh& m);

// A cell order that groups cells of the same shape class (edge directions from the first
// vertex, rounded), stable within a class: new cell i = old cell perm[i]. On a structured
// simplex subdivision (the Kuhn cube) a class is one cell type in grid order: the cells of
// a wave then find their neighbours at a common offset through the same local nodes, so
// the face gathers are contiguous. mode 0: over the whole mesh; 1 (windowed): within runs of
// 64 K cells (K classes, at most 32), which keeps each wave to one class and a cell's
// neighbours within a few runs of it; 2 (tiled): within space tiles holding about 64 cells of
// each class, tiles in lexicographic order, so most neighbours share a cell's tile.
std::vector<int64_t> shape_class_order(const HostMesh& m, int mode);
// The same mesh with cells renumbered, new cell i = old cell perm[i].
HostMesh renumber_cells(const HostMesh& m, const std::vector<int64_t>& perm);

// d xi_a / d x_x of every cell, cells x d x d.
void mesh_inverse_jacobians(const HostMesh& m, double* G);

struct RefElement;
// Nodal coordinates of an order-N element set, dim x cells x Np.
void mesh_node_coords(const HostMesh& m, const RefElement& re, double* xyz);

}  // namespace insp
