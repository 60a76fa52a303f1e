// Structured grid, compartments, DOF layout and BCSR pattern.
//
// Mirrors what the reference sets up once per run:
//   * the structured grid of `grid.cells` / `grid.extensions` / `grid.origin`,
//     refined `grid.refinement_level` times; Kuhn simplices in 2D/3D as
//     createSimplexGrid makes them (grid/make_multi_domain_grid.hh:76-91), or
//     Q1 boxes with `grid.cell_type = cube`;
//   * compartments by element centre (make_multi_domain_grid.hh:118-194);
//   * per compartment, the vertices its elements touch, numbered in increasing
//     vertex order, with the compartment's species interleaved per vertex
//     ("EntityGrouping", model_single_compartment_traits.hh:26-27), and
//     compartments in INI order (model_multi_compartment.impl.hh:63-83);
//   * the Jacobian sparsity (local_operator.hh:276-399,
//     model/make_step_operator.hh:295-401) as one BCSR matrix whose blocks
//     are B x B with B the largest species count of any compartment.
//
// Device layout ("slot" = one vertex of one compartment):
//   row_ptr[nslot+1], col[nnzb]      int32 BCSR structure, columns sorted
//   val[nnzb][B][B]                  row-major blocks
//   own_start[nslot], own_mask[nslot] position of the first same-compartment
//                                    column and a 3^d-bit stencil mask, so an
//                                    element finds its block (a,c) as
//                                    own_start[a] + popcount(mask & ((1<<o)-1))
//                                    with o the stencil offset of c from a.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "expr.hh"

namespace dcopasi {

struct Grid {
  int dim = 2;
  bool cube = false;  // Q1 boxes, else Kuhn simplices (lines in 1D)
  std::array<long, 3> n{1, 1, 1};  // cells per axis (1 beyond dim)
  std::array<double, 3> origin{0, 0, 0}, h{1, 1, 1};
  long nvert = 0, ncell = 0;
  int nclass = 1;   // simplices per cell (d!) or 1
  int nvpe = 2;     // vertices per element
  int ncorner = 2;  // vertices per cell (2^d)
  // corner[cls][a]: cell-corner index (bit k = +1 along axis k) of local vertex a
  std::vector<std::array<int, 8>> corner;
  // unstructured simplex mesh read from `grid.path`: every cell is one
  // element (one class, corners = its vertices in file order)
  bool unstructured = false;
  std::vector<double> X;        // 3 coordinates per vertex
  std::vector<int> conn;        // nvpe vertices per cell
  std::vector<double> gmsh_id;  // physical entity of each cell

  long vertex(long i, long j, long k) const { return i + (n[0] + 1) * (j + (n[1] + 1) * k); }
  long cell_vertex(long cell, int c) const;  // global vertex of a cell corner
  void cell_ijk(long cell, long ijk[3]) const;
  void vertex_pos(long v, double x[3]) const;
};

struct Facet {
  // one side: compartment, slots of the facet's vertices (nvf), and for
  // boundary facets out == in.
  int comp_in = -1, comp_out = -1;
  bool boundary = false;
  std::vector<int> slot_in, slot_out;
  // per quadrature point: weight*|J_F|, position, unit normal (outward of in)
  std::vector<double> qw, qx, qn;
  std::vector<double> phi;  // [q][f] facet basis values
  double volume = 0.0;
};

struct Compartment {
  std::string name;
  std::vector<int> species;  // global species indices, INI order
  int b = 0;                 // species count
  long nslot = 0, slot_off = 0, dof_off = 0;
  std::vector<int> g2l;       // vertex -> local node or -1
  std::vector<long> nodes;    // local node -> vertex
  // assembly work list: cells touched by this compartment, sorted by colour
  std::vector<int> cells;
  std::vector<uint8_t> cell_mask;       // which Kuhn classes of the cell belong here
  std::vector<int> cell_slots;          // [cell][corner] slot or -1
  std::vector<int> color_ptr;           // colour ranges into cells
  std::vector<int> cell_pos;            // unstructured: [cell][a][c] BCSR block of (a, c)
};

struct Layout {
  Grid grid;
  std::vector<Compartment> comps;
  int B = 1;                 // block size
  long nslot = 0, ndof = 0;  // ndof: packed DOF count (no padding)
  std::vector<int> slot_comp;
  std::vector<int> row_ptr, col;
  std::vector<int> own_start;
  std::vector<uint32_t> own_mask;
  std::vector<int> diag_pos;  // position of the diagonal block per row
  std::vector<Facet> facets;
  std::vector<int> facet_color_ptr;  // facets sorted by colour
  long nnzb() const { return static_cast<long>(col.size()); }
};

struct SpeciesDecl {
  std::string name;
  int comp;
};

// Build the grid from `grid.*` keys.
Grid make_grid(int dim, const std::vector<double>& cells, const std::vector<double>& ext,
               const std::vector<double>& origin, int refine, bool cube);

// Read a gmsh file (format 2, ASCII), as dune-grid's GmshReader does for the
// reference (make_multi_domain_grid.hh:56-63): the elements of the grid
// dimension (triangles in 2D, tetrahedra in 3D, lines in 1D) become cells,
// their first tag (the physical entity) becomes the cell datum `gmsh_id`, and
// the vertices they use are numbered in increasing node number.
Grid read_gmsh(const std::string& path, int dim);

// `levels` rounds of uniform red refinement of an unstructured simplex grid
// (grid.refinement_level on a grid read from a file; the reference calls
// globalRefine on its UGGrid, make_multi_domain_grid.hh:96-100).  Every edge
// gets its midpoint; a line splits in 2, a triangle in 4 (three corners and
// the middle one), a tetrahedron in 8 (four corners and the inner octahedron
// cut along its shortest diagonal, the first of m01-m23, m02-m13, m03-m12 on a
// tie).  Children keep the parent's gmsh_id.  Vertices keep their indices and
// the midpoints follow in the order they are first met (cells in order, edges
// 01 02 03 12 13 23); each cell's children replace it in place.
void refine_simplices(Grid& g, int levels);

// Assign elements to compartments by evaluating each compartment's
// expression at element centres; returns per element (cell*nclass+cls) the
// compartment or -1.  Overlapping compartments are rejected.
std::vector<int8_t> mark_elements(const Grid& g, const std::vector<expr::Ex>& comp_expr);

// Everything after marking: slots, work lists, colours, pattern, facets.
// `need_facets[c]` tells whether species of compartment c have outflow terms
// (boundary facets of c) and `interface` whether any interface term exists.
Layout make_layout(const Grid& g, const std::vector<int8_t>& elem_comp, std::vector<Compartment> comps,
                   bool boundary_facets, bool interface_facets);

// quadrature on reference elements (order 2, as local_operator.hh:128-139)
struct Rule {
  std::vector<std::array<double, 3>> x;
  std::vector<double> w;
};
Rule element_rule(int dim, bool cube);
// order-4 rule of the transform/reduce evaluation (reduce.hh:167-169): Gauss-
// Legendre with 3 points per axis on boxes, the 6-point degree-4 rule on
// triangles and the 11-point degree-4 rule on tetrahedra
Rule reduce_rule(int dim, bool cube);
Rule facet_rule(int dim, bool cube);

}  // namespace dcopasi
