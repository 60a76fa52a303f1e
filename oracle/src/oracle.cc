// CPU oracle for the DuneCopasi diffusion-reaction per-time-step path.
//
// TEST INFRASTRUCTURE.  This library is the parity checker and the timed CPU
// baseline; only the tests, the smoke entry and the benchmark's CPU leg load
// it.  The product never links or calls it.
//
// It restates, in plain sequential/OpenMP C++, what the reference computes:
//   * structured grids and compartments by cell centre
//       dune/copasi/grid/make_multi_domain_grid.hh:76-194
//   * node-grouped DOF layout per compartment, compartments in INI order
//       model/diffusion_reaction/model_single_compartment_traits.hh:26-27,
//       model_multi_compartment.impl.hh:63-83
//   * the volume residual and analytical Jacobian of the local operator
//       model/diffusion_reaction/local_operator.hh:417-491, 541-707
//     (with the exact derivative where the reference's advection and
//      nonlinear cross-diffusion Jacobians deviate from it, :648-655, :688-699)
//   * the numerical (forward difference) Jacobian  local_operator.hh:713-765
//   * transmission / boundary outflow terms        local_operator.hh:777-971,
//       1398-1415, with exact derivatives for all four coupling blocks
//   * which terms belong to mass vs stiffness form local_equations.hh:322-355
//   * Newton / linear defect correction / RK       model/make_step_operator.hh:183-443
//   * ILU0 (istl SeqILU, n=0), Jacobi (SeqJac), BiCGSTAB (istl)
//       solver/istl/factory/{iterative,preconditioner}.hh
//   * per-compartment mass sums                    model/diffusion_reaction/reduce.hh:160-185
//
// Matrices are scalar CSR over the flat DOF vector (the reference's
// BCRSMatrix<BCRSMatrix<double>> with a 1x1 Runge-Kutta outer block,
// make_step_operator.hh:46-53,298-303), not the product's blocked format.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <limits>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "expr.hh"
#include "ini.hh"

namespace orc {

// ---------------------------------------------------------------- quadrature
// Reference-element rules of order 2 (what the reference requests: 2 * order,
// local_operator.hh:128-139).  Simplices: the 3-point interior triangle rule
// and the 4-point tetrahedron rule; cubes and lines: 2-point Gauss-Legendre per
// axis.  All are exact for the mass, stiffness and linear reaction integrands
// of P1/Q1.
struct QRule {
  std::vector<std::array<double, 3>> x;
  std::vector<double> w;
};

static QRule gauss_tensor(int d) {
  const double g0 = 0.5 - 0.5 / std::sqrt(3.0), g1 = 0.5 + 0.5 / std::sqrt(3.0);
  const double g[2] = {g0, g1};
  QRule r;
  int n = 1 << d;
  for (int i = 0; i < n; ++i) {
    std::array<double, 3> p{0, 0, 0};
    for (int k = 0; k < d; ++k) p[k] = g[(i >> k) & 1];
    r.x.push_back(p);
    r.w.push_back(1.0 / n);
  }
  return r;
}

static QRule simplex_rule(int d) {
  QRule r;
  if (d == 0) {
    r.x.push_back({0, 0, 0});
    r.w.push_back(1.0);
  } else if (d == 1) {
    r = gauss_tensor(1);
  } else if (d == 2) {
    const double a = 1.0 / 6.0, b = 2.0 / 3.0;
    r.x = {{a, a, 0}, {b, a, 0}, {a, b, 0}};
    r.w = {1.0 / 6, 1.0 / 6, 1.0 / 6};
  } else {
    const double a = (5.0 + 3.0 * std::sqrt(5.0)) / 20.0, b = (5.0 - std::sqrt(5.0)) / 20.0;
    r.x = {{b, b, b}, {a, b, b}, {b, a, b}, {b, b, a}};
    r.w = {1.0 / 24, 1.0 / 24, 1.0 / 24, 1.0 / 24};
  }
  return r;
}

static QRule cube_rule(int d) {
  if (d == 0) return simplex_rule(0);
  return gauss_tensor(d);
}

// ---------------------------------------------------------------- mesh
struct Mesh {
  int dim = 2;
  bool cube = false;  // Q1 cells when true, Kuhn-split P1 simplices otherwise
  int nvpe = 3;       // vertices per element
  std::array<long, 3> nc{1, 1, 1};
  std::array<double, 3> origin{0, 0, 0}, ext{1, 1, 1};
  long nvert = 0, nelem = 0;
  std::vector<double> X;    // 3 * nvert
  std::vector<long> conn;   // nvpe * nelem
  std::vector<double> phys; // gmsh physical entity per element (grids read from files)
};

// gmsh format 2 (ASCII), read as dune-grid's GmshReader feeds the reference
// (make_multi_domain_grid.hh:56-63): elements of the grid dimension are the
// cells (lower-dimensional ones are boundary segments), the first tag is the
// element's physical entity (cell datum gmsh_id), and the nodes the cells use
// are numbered in increasing node number.
static void read_gmsh(Mesh& m, const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("oracle: cannot open '" + path + "'");
  const int want = m.dim == 1 ? 1 : m.dim == 2 ? 2 : 4;
  const int nvpe = m.dim + 1;
  std::map<long, std::array<double, 3>> node;
  std::vector<std::vector<long>> cells;
  std::vector<double> phys;
  std::string w;
  while (in >> w) {
    if (w == "$Nodes") {
      long n;
      in >> n;
      for (long i = 0; i < n; ++i) {
        long id;
        std::array<double, 3> x;
        in >> id >> x[0] >> x[1] >> x[2];
        if (!in) throw std::runtime_error("oracle: malformed $Nodes in '" + path + "'");
        node[id] = x;
      }
    } else if (w == "$Elements") {
      long n;
      in >> n;
      for (long i = 0; i < n; ++i) {
        long id;
        int type, nt;
        in >> id >> type >> nt;
        std::vector<long> tag(nt);
        for (auto& t : tag) in >> t;
        int nv = type == 15 ? 1 : type == 1 ? 2 : type == 2 ? 3 : type == 4 ? 4 : -1;
        if (nv < 0) throw std::runtime_error("oracle: gmsh element type " + std::to_string(type) + " not restated");
        std::vector<long> v(nv);
        for (auto& x : v) in >> x;
        if (!in) throw std::runtime_error("oracle: malformed $Elements in '" + path + "'");
        if (type == want) {
          cells.push_back(v);
          phys.push_back(nt ? (double)tag[0] : 0.0);
        }
      }
    }
  }
  if (cells.empty()) throw std::runtime_error("oracle: no cells of dimension " + std::to_string(m.dim) + " in '" + path + "'");
  std::vector<long> ids;
  for (auto& c : cells) ids.insert(ids.end(), c.begin(), c.end());
  std::sort(ids.begin(), ids.end());
  ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
  std::map<long, long> idx;
  m.cube = false;
  m.nvpe = nvpe;
  m.nvert = (long)ids.size();
  m.nelem = (long)cells.size();
  m.X.assign(3 * m.nvert, 0.0);
  for (long i = 0; i < m.nvert; ++i) {
    idx[ids[i]] = i;
    for (int k = 0; k < m.dim; ++k) m.X[3 * i + k] = node.at(ids[i])[k];
  }
  m.conn.clear();
  for (auto& c : cells)
    for (int a = 0; a < nvpe; ++a) m.conn.push_back(idx[c[a]]);
  m.phys = phys;
}

// Uniform red refinement of a mesh read from a file, `lvl` times (the
// reference's UGGrid globalRefine, make_multi_domain_grid.hh:96-100): edge
// midpoints appended in first-met order (elements in order, local edges
// 01 02 03 12 13 23), children in place of their parent: lines 2, triangles
// 4 (corner children, then the middle), tetrahedra 8 (corner children, then
// the octahedron split along its shortest diagonal, first of m01-m23,
// m02-m13, m03-m12 on ties); children keep the parent's physical entity.
static void refine_file_mesh(Mesh& m, int lvl) {
  const int d = m.dim, nv = m.nvpe;
  std::vector<std::pair<int, int>> edges;
  for (int a = 0; a < nv; ++a)
    for (int b = a + 1; b < nv; ++b) edges.push_back({a, b});
  for (int l = 0; l < lvl; ++l) {
    std::map<std::pair<long, long>, long> midpoint;
    std::vector<long> conn;
    std::vector<double> phys;
    for (long e = 0; e < m.nelem; ++e) {
      std::vector<long> v(m.conn.begin() + e * nv, m.conn.begin() + (e + 1) * nv);
      // M[a][b]: vertex a for a == b, else the midpoint of edge ab
      long M[4][4];
      for (int a = 0; a < nv; ++a) M[a][a] = v[a];
      for (auto [a, b] : edges) {
        std::pair<long, long> key{std::min(v[a], v[b]), std::max(v[a], v[b])};
        auto it = midpoint.find(key);
        long id;
        if (it != midpoint.end()) {
          id = it->second;
        } else {
          id = (long)(m.X.size() / 3);
          for (int k = 0; k < 3; ++k) m.X.push_back((m.X[3 * v[a] + k] + m.X[3 * v[b] + k]) / 2.0);
          midpoint[key] = id;
        }
        M[a][b] = M[b][a] = id;
      }
      std::vector<std::vector<long>> kids;
      for (int a = 0; a < nv; ++a) {
        // corner child at vertex a: a and the midpoints of its edges, in the parent's order
        std::vector<long> c(nv);
        for (int b = 0; b < nv; ++b) c[b] = M[a][b];
        kids.push_back(c);
      }
      if (d == 2) kids.push_back({M[0][1], M[1][2], M[0][2]});
      if (d == 3) {
        auto dist2 = [&](long p, long q) {
          double s = 0;
          for (int k = 0; k < 3; ++k) s += (m.X[3 * p + k] - m.X[3 * q + k]) * (m.X[3 * p + k] - m.X[3 * q + k]);
          return s;
        };
        // diagonals as vertex pairs (ab, cd) with {a, b, c, d} = {0, 1, 2, 3}
        const int dg[3][4] = {{0, 1, 2, 3}, {0, 2, 1, 3}, {0, 3, 1, 2}};
        int k0 = 0;
        for (int k = 1; k < 3; ++k)
          if (dist2(M[dg[k][0]][dg[k][1]], M[dg[k][2]][dg[k][3]]) < dist2(M[dg[k0][0]][dg[k0][1]], M[dg[k0][2]][dg[k0][3]]))
            k0 = k;
        const int a = dg[k0][0], b = dg[k0][1], c = dg[k0][2], e2 = dg[k0][3];
        const long p = M[a][b], q = M[c][e2];
        // the other four midpoints, each sharing one index with its neighbour
        const long ring[4] = {M[a][c], M[a][e2], M[b][e2], M[b][c]};
        for (int k = 0; k < 4; ++k) kids.push_back({p, q, ring[k], ring[(k + 1) % 4]});
      }
      for (auto& c : kids) {
        conn.insert(conn.end(), c.begin(), c.end());
        phys.push_back(m.phys[e]);
      }
    }
    m.conn.swap(conn);
    m.phys.swap(phys);
    m.nvert = (long)(m.X.size() / 3);
    m.nelem = (long)m.phys.size();
  }
}

static long vid(const Mesh& m, long i, long j, long k) {
  return i + (m.nc[0] + 1) * (j + (m.nc[1] + 1) * k);
}

// Structured grid on [origin, origin+ext].  dim 1 always uses line cells (the
// reference's createCubeGrid); dim >= 2 uses Kuhn simplices (createSimplexGrid,
// make_multi_domain_grid.hh:87-89) unless cube cells are requested.  The Kuhn
// split follows the main diagonal of each cell: one simplex per permutation of
// the axes, vertices along the monotone path corner -> opposite corner.
static void build_mesh(Mesh& m) {
  int d = m.dim;
  for (int k = d; k < 3; ++k) m.nc[k] = 0;
  long nv = 1;
  for (int k = 0; k < d; ++k) nv *= (m.nc[k] + 1);
  m.nvert = nv;
  m.X.assign(3 * nv, 0.0);
  for (long k = 0; k <= m.nc[2]; ++k)
    for (long j = 0; j <= m.nc[1]; ++j)
      for (long i = 0; i <= m.nc[0]; ++i) {
        long v = vid(m, i, j, k);
        long idx[3] = {i, j, k};
        for (int a = 0; a < d; ++a) m.X[3 * v + a] = m.origin[a] + m.ext[a] * double(idx[a]) / double(m.nc[a]);
      }
  if (d == 1) m.cube = true;
  std::vector<std::vector<int>> perms;
  if (!m.cube) {
    std::vector<int> p(d);
    std::iota(p.begin(), p.end(), 0);
    do perms.push_back(p); while (std::next_permutation(p.begin(), p.end()));
  }
  m.nvpe = m.cube ? (1 << d) : d + 1;
  long ncell = 1;
  for (int k = 0; k < d; ++k) ncell *= m.nc[k];
  m.nelem = m.cube ? ncell : ncell * (long)perms.size();
  m.conn.resize((size_t)m.nelem * m.nvpe);
  long e = 0;
  for (long k = 0; k < std::max<long>(1, m.nc[2]); ++k)
    for (long j = 0; j < std::max<long>(1, m.nc[1]); ++j)
      for (long i = 0; i < m.nc[0]; ++i) {
        long base[3] = {i, j, k};
        if (m.cube) {
          for (int v = 0; v < m.nvpe; ++v) {
            long c[3] = {base[0], base[1], base[2]};
            for (int a = 0; a < d; ++a) c[a] += (v >> a) & 1;
            m.conn[e * m.nvpe + v] = vid(m, c[0], c[1], c[2]);
          }
          ++e;
        } else {
          for (auto& p : perms) {
            long c[3] = {base[0], base[1], base[2]};
            m.conn[e * m.nvpe + 0] = vid(m, c[0], c[1], c[2]);
            for (int s = 0; s < d; ++s) {
              c[p[s]] += 1;
              m.conn[e * m.nvpe + s + 1] = vid(m, c[0], c[1], c[2]);
            }
            ++e;
          }
        }
      }
}

// Affine element geometry x = x0 + J xi (simplices; axis-aligned boxes).
struct Geo {
  double x0[3] = {0, 0, 0};
  double J[3][3] = {{0}};
  double Jinv[3][3] = {{0}};
  double det = 0;  // |det J|
};

static void invert(int d, const double (*A)[3], double (*B)[3], double& det) {
  if (d == 1) {
    det = A[0][0];
    B[0][0] = 1.0 / A[0][0];
  } else if (d == 2) {
    det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    B[0][0] = A[1][1] / det;
    B[0][1] = -A[0][1] / det;
    B[1][0] = -A[1][0] / det;
    B[1][1] = A[0][0] / det;
  } else {
    det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
          A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    B[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
    B[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
    B[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
    B[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
    B[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
    B[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
    B[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
    B[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
    B[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
  }
  det = std::fabs(det);
}

static Geo element_geo(const Mesh& m, long e) {
  Geo g;
  int d = m.dim;
  const long* c = &m.conn[e * m.nvpe];
  for (int a = 0; a < 3; ++a) g.x0[a] = m.X[3 * c[0] + a];
  for (int k = 0; k < d; ++k) {
    long v = m.cube ? c[1 << k] : c[k + 1];
    for (int a = 0; a < d; ++a) g.J[a][k] = m.X[3 * v + a] - g.x0[a];
  }
  invert(d, g.J, g.Jinv, g.det);
  return g;
}

// shape functions on the reference element
static void shape(const Mesh& m, const double* xi, double* phi, double (*dphi)[3]) {
  int d = m.dim;
  if (m.cube) {
    for (int v = 0; v < m.nvpe; ++v) {
      double p = 1.0;
      double g[3] = {1, 1, 1};
      for (int a = 0; a < d; ++a) {
        bool b = (v >> a) & 1;
        double f = b ? xi[a] : 1.0 - xi[a];
        double df = b ? 1.0 : -1.0;
        p *= f;
        for (int c = 0; c < d; ++c) g[c] *= (c == a) ? df : f;
      }
      phi[v] = p;
      for (int c = 0; c < 3; ++c) dphi[v][c] = c < d ? g[c] : 0.0;
    }
  } else {
    double s = 0;
    for (int a = 0; a < d; ++a) s += xi[a];
    phi[0] = 1.0 - s;
    for (int c = 0; c < 3; ++c) dphi[0][c] = c < d ? -1.0 : 0.0;
    for (int v = 1; v <= d; ++v) {
      phi[v] = xi[v - 1];
      for (int c = 0; c < 3; ++c) dphi[v][c] = (c == v - 1) ? 1.0 : 0.0;
    }
  }
}

// local vertex lists of the faces of a reference element
static std::vector<std::vector<int>> faces(const Mesh& m) {
  std::vector<std::vector<int>> f;
  int d = m.dim;
  if (m.cube) {
    for (int a = 0; a < d; ++a)
      for (int s = 0; s < 2; ++s) {
        std::vector<int> l;
        for (int v = 0; v < m.nvpe; ++v)
          if (((v >> a) & 1) == s) l.push_back(v);
        f.push_back(l);
      }
  } else {
    for (int k = 0; k <= d; ++k) {
      std::vector<int> l;
      for (int v = 0; v <= d; ++v)
        if (v != k) l.push_back(v);
      f.push_back(l);
    }
  }
  return f;
}

// ---------------------------------------------------------------- model
enum Slot { V_TIME = 0, V_POS = 1, V_NORMAL = 4, V_EVOL = 7, V_INVOL = 8, V_INBND = 9, V_INSKEL = 10, V_IFAC = 11, V_SPEC = 12 };

struct Jac {
  int wrt;  // global species index
  Expr e;
};
struct Diff {
  int wrt;
  bool tensor = false;
  Expr d[3][3];  // scalar in d[0][0]
  struct DJac {
    int wrt;
    bool tensor;
    Expr d[3][3];
  };
  std::vector<DJac> jac;
};
struct Outflow {
  int to_comp = -1;
  Expr e;
  std::vector<Jac> jac;
};
struct Species {
  std::string name;
  int comp = -1, local = -1;
  Expr storage, reaction, velocity[3];
  std::vector<Jac> storage_jac, reaction_jac;
  std::vector<std::array<Expr, 3>> vel_jac;
  std::vector<int> vel_jac_wrt;
  std::vector<Diff> diff;
  std::vector<Outflow> outflow;
};
struct Compartment {
  std::string name;
  std::vector<int> species;
  std::vector<long> elems, nodes, g2l;
  long dof_off = 0, slot_off = 0;
  int b = 0;
};
// facet contribution record: element, local face, neighbour element or -1
struct Facet {
  long e_in, e_out;
  int f_in;
  int c_in, c_out;  // compartments (c_out == c_in for boundary facets)
};

struct RK {
  int stages;
  std::vector<std::vector<double>> a, b;
  std::vector<double> d;
};

static RK rk_table(const std::string& type) {
  // PDELab one-step form of each scheme (the tables of
  // dune/pdelab/instationary/onestepparameter.hh the reference selects in
  // model/make_step_operator.hh:408-443): stage r solves
  //   sum_j a[r][j] M(u_j) + dt b[r][j] A(u_j, t + d[j] dt) = 0
  RK r;
  const double s2 = std::sqrt(2.0);
  if (type == "ExplicitEuler") {
    r.stages = 1;
    r.a = {{-1.0, 1.0}};
    r.b = {{1.0, 0.0}};
    r.d = {0.0, 1.0};
  } else if (type == "ImplicitEuler") {
    r.stages = 1;
    r.a = {{-1.0, 1.0}};
    r.b = {{0.0, 1.0}};
    r.d = {0.0, 1.0};
  } else if (type == "Heun") {
    r.stages = 2;
    r.a = {{-1.0, 1.0, 0.0}, {-0.5, -0.5, 1.0}};
    r.b = {{1.0, 0.0, 0.0}, {0.0, 0.5, 0.0}};
    r.d = {0.0, 1.0, 1.0};
  } else if (type == "Shu3") {
    r.stages = 3;
    r.a = {{-1.0, 1.0, 0.0, 0.0}, {-3.0 / 4.0, -1.0 / 4.0, 1.0, 0.0}, {-1.0 / 3.0, 0.0, -2.0 / 3.0, 1.0}};
    r.b = {{1.0, 0.0, 0.0, 0.0}, {0.0, 1.0 / 4.0, 0.0, 0.0}, {0.0, 0.0, 2.0 / 3.0, 0.0}};
    r.d = {0.0, 1.0, 0.5, 1.0};
  } else if (type == "RungeKutta4") {
    r.stages = 4;
    r.a.assign(4, std::vector<double>(5, 0.0));
    r.b.assign(4, std::vector<double>(5, 0.0));
    for (int k = 0; k < 4; ++k) {
      r.a[k][0] = -1.0;
      r.a[k][k + 1] = 1.0;
    }
    r.b[0][0] = 0.5;
    r.b[1][1] = 0.5;
    r.b[2][2] = 1.0;
    r.b[3] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0, 0.0};
    r.d = {0.0, 0.5, 0.5, 1.0, 1.0};
  } else if (type == "Alexander2") {
    // Alexander (1977) 2-stage L-stable SDIRK, alpha = 1 - sqrt(2)/2
    double al = 1.0 - s2 / 2.0;
    r.stages = 2;
    r.a = {{-1.0, 1.0, 0.0}, {-1.0, 0.0, 1.0}};
    r.b = {{0.0, al, 0.0}, {0.0, 1.0 - al, al}};
    r.d = {0.0, al, 1.0};
  } else if (type == "FractionalStepTheta") {
    // theta-steps theta, 1 - 2 theta, theta; implicit weights alpha, beta, alpha
    const double th = 1.0 - s2 / 2.0, thp = 1.0 - 2.0 * th, al = (1.0 - 2.0 * th) / (1.0 - th), be = 1.0 - al;
    r.stages = 3;
    r.a = {{-1.0, 1.0, 0.0, 0.0}, {0.0, -1.0, 1.0, 0.0}, {0.0, 0.0, -1.0, 1.0}};
    r.b = {{be * th, al * th, 0.0, 0.0}, {0.0, al * thp, be * thp, 0.0}, {0.0, 0.0, be * th, al * th}};
    r.d = {0.0, th, 1.0 - th, 1.0};
  } else if (type == "Alexander3") {
    // 3-stage L-stable SDIRK of order 3; alpha solves
    // 6 x^3 - 18 x^2 + 9 x - 1 = 0 on (0.4, 0.5), here by bisection
    double lo = 0.4, hi = 0.5;
    for (int it = 0; it < 200; ++it) {
      const double mid = 0.5 * (lo + hi);
      const double f = ((6.0 * mid - 18.0) * mid + 9.0) * mid - 1.0;
      (f > 0.0 ? lo : hi) = mid;
    }
    const double al = 0.5 * (lo + hi), tau = 0.5 * (1.0 + al);
    const double b1 = -(6.0 * al * al - 16.0 * al + 1.0) / 4.0, b2 = (6.0 * al * al - 20.0 * al + 5.0) / 4.0;
    r.stages = 3;
    r.a = {{-1.0, 1.0, 0.0, 0.0}, {-1.0, 0.0, 1.0, 0.0}, {-1.0, 0.0, 0.0, 1.0}};
    r.b = {{0.0, al, 0.0, 0.0}, {0.0, tau - al, al, 0.0}, {0.0, b1, b2, al}};
    r.d = {0.0, al, tau, 1.0};
  } else {
    throw std::runtime_error("oracle: time stepping type '" + type + "' not supported");
  }
  return r;
}

struct Model {
  Ini ini;
  Scope scope;
  Mesh mesh;
  std::vector<Compartment> comps;
  std::vector<Species> spec;
  std::vector<int> elem_comp;  // compartment of each element or -1
  std::vector<Facet> facets;
  bool has_outflow = false;
  bool numerical_jac = false;
  double fd_eps = 1e-7;
  bool is_linear = false;
  long ndof = 0, nslot = 0;
  // slot adjacency (sorted) and scalar CSR
  std::vector<long> adj_ptr;
  std::vector<long> adj;
  std::vector<long> adj_off;  // cumulative species count of neighbours before entry
  std::vector<long> row_ptr;  // ndof + 1
  std::vector<int> col;
  std::vector<long> slot_dof;  // first dof of a slot
  std::vector<int> slot_b;     // species count of a slot
  std::vector<std::vector<long>> colors;  // element colors per compartment-major list
  // state
  std::vector<double> u;
  double time = 0;
  RK rk;
  // solver settings
  std::string lin_type = "BiCGSTAB", precond = "ILU";
  double lin_rtol = 1e-4;
  int lin_maxit = 500;
  double nl_rtol = 1e-4, nl_atol = 0.0;
  int nl_min = 0, nl_max = 40;
  std::string nl_norm = "l_2";
  bool fixed_lin_tol = false;
  double min_lin_red = 0.1, lin_threshold = 0.0;
  // last statistics
  int last_newton = 0, last_linear = 0;
  double last_assembly_s = 0, last_solve_s = 0;
};

static int nthreads_req = 0;

static void build_species(Model& M) {
  const Ini& I = M.ini;
  // compartments in INI order; each gets a cell-centre expression
  auto cnames = I.subkeys("compartments");
  // an explicit [compartments] section may carry plain "name.expression" keys
  std::map<std::string, int> cid;
  for (auto& c : cnames) {
    Compartment C;
    C.name = c;
    cid[c] = (int)M.comps.size();
    M.comps.push_back(C);
  }
  auto snames = I.subkeys("model.scalar_field");
  for (auto& s : snames) {
    std::string cname = I.get("model.scalar_field." + s + ".compartment");
    if (cname.empty()) continue;
    auto it = cid.find(cname);
    if (it == cid.end()) throw std::runtime_error("oracle: species '" + s + "' in unknown compartment '" + cname + "'");
    Species S;
    S.name = s;
    S.comp = it->second;
    S.local = (int)M.comps[S.comp].species.size();
    M.comps[S.comp].species.push_back((int)M.spec.size());
    M.spec.push_back(S);
  }
  // compartments without species take no DOFs (the reference builds them empty)
  for (auto& C : M.comps) C.b = (int)C.species.size();
}

// parser_context.<name>.{domain,range} of an interpolation function
// (context.cc:72-97), checked as the reference checks them
static std::shared_ptr<const Table> read_table(const Ini& I, const std::string& p) {
  auto t = std::make_shared<Table>();
  t->x = I.get_doubles(p + ".domain");
  t->y = I.get_doubles(p + ".range");
  if (!std::is_sorted(t->x.begin(), t->x.end())) throw std::runtime_error("The interpolation domain must be sorted");
  if (t->x.size() < 2 || t->x.size() != t->y.size())
    throw std::runtime_error("Interpolation range and domain must have at least two points and be the same size");
  return t;
}

static void build_scope(Model& M) {
  const Ini& I = M.ini;
  Scope& sc = M.scope;
  int d = M.mesh.dim;
  const char* ax = "xyz";
  sc.vars["time"] = V_TIME;
  sc.vars["entity_volume"] = V_EVOL;
  sc.vars["in_volume"] = V_INVOL;
  sc.vars["in_boundary"] = V_INBND;
  sc.vars["in_skeleton"] = V_INSKEL;
  sc.vars["integration_factor"] = V_IFAC;
  for (int a = 0; a < 3; ++a) {
    std::string p = std::string("position_") + ax[a];
    std::string n = std::string("normal_") + ax[a];
    if (a < d) {
      sc.vars[p] = V_POS + a;
      sc.vars[n] = V_NORMAL + a;
    } else {
      sc.consts[p] = 0.0;
      sc.consts[n] = 0.0;
    }
  }
  for (size_t s = 0; s < M.spec.size(); ++s) {
    sc.vars[M.spec[s].name] = V_SPEC + 4 * (int)s;
    for (int a = 0; a < 3; ++a) {
      std::string g = "grad_" + M.spec[s].name + "_" + ax[a];
      if (a < d) sc.vars[g] = V_SPEC + 4 * (int)s + 1 + a;
      else sc.consts[g] = 0.0;
    }
  }
  // parser_context: constants and inline functions
  for (auto& k : I.subkeys("parser_context")) {
    std::string p = "parser_context." + k;
    std::string type = I.get(p + ".type", "");
    if (type == "constant") {
      sc.consts[k] = std::stod(I.get(p + ".value"));
    } else if (type == "function") {
      std::string e = I.get(p + ".expression");
      auto colon = e.find(':');
      if (colon == std::string::npos) throw std::runtime_error("oracle: function '" + k + "' needs 'args: body'");
      UserFunction f;
      std::string args = e.substr(0, colon);
      size_t st = 0;
      while (st <= args.size()) {
        size_t c = args.find(',', st);
        std::string a = trim(args.substr(st, c == std::string::npos ? std::string::npos : c - st));
        if (!a.empty()) f.args.push_back(a);
        if (c == std::string::npos) break;
        st = c + 1;
      }
      f.body = e.substr(colon + 1);
      sc.funcs[k] = f;
    } else if (type == "interpolation") {
      sc.tables[k] = read_table(I, p);
    } else if (!type.empty()) {
      throw std::runtime_error("oracle: parser_context type '" + type + "' not supported");
    }
  }
}

static Expr cexpr(Model& M, const std::string& key) { return compile(M.ini.get(key, ""), M.scope); }

static std::vector<Jac> jac_list(Model& M, const std::string& prefix) {
  std::vector<Jac> r;
  for (size_t s = 0; s < M.spec.size(); ++s) {
    std::string k = prefix + ".jacobian." + M.spec[s].name + ".expression";
    if (!M.ini.has(k)) continue;
    Jac j{(int)s, cexpr(M, k)};
    if (!j.e.null()) r.push_back(j);
  }
  return r;
}

static bool read_diff(Model& M, const std::string& p, bool& tensor, Expr (&d)[3][3]) {
  std::string type = M.ini.get(p + ".type", "scalar");
  const char* ax = "xyz";
  bool any = false;
  if (type == "scalar") {
    tensor = false;
    d[0][0] = cexpr(M, p + ".expression");
    any = !d[0][0].null();
  } else if (type == "tensor") {
    tensor = true;
    for (int i = 0; i < M.mesh.dim; ++i)
      for (int j = 0; j < M.mesh.dim; ++j) {
        d[i][j] = cexpr(M, p + "." + ax[i] + ax[j] + ".expression");
        any |= !d[i][j].null();
      }
  } else {
    throw std::runtime_error("oracle: unknown diffusion type '" + type + "'");
  }
  return any;
}

static void build_equations(Model& M) {
  const char* ax = "xyz";
  for (auto& S : M.spec) {
    std::string p = "model.scalar_field." + S.name;
    S.storage = cexpr(M, p + ".storage.expression");
    if (!S.storage.null()) S.storage_jac = jac_list(M, p + ".storage");
    S.reaction = cexpr(M, p + ".reaction.expression");
    if (!S.reaction.null()) S.reaction_jac = jac_list(M, p + ".reaction");
    bool anyv = false;
    for (int a = 0; a < M.mesh.dim; ++a) {
      S.velocity[a] = cexpr(M, p + ".velocity." + ax[a] + ".expression");
      anyv |= !S.velocity[a].null();
    }
    if (anyv) {
      for (size_t w = 0; w < M.spec.size(); ++w) {
        std::array<Expr, 3> v;
        bool any = false;
        for (int a = 0; a < M.mesh.dim; ++a) {
          v[a] = cexpr(M, p + ".velocity.jacobian." + M.spec[w].name + "." + ax[a] + ".expression");
          any |= !v[a].null();
        }
        if (any) {
          S.vel_jac.push_back(v);
          S.vel_jac_wrt.push_back((int)w);
        }
      }
    } else {
      for (auto& v : S.velocity) v = Expr{};
    }
    // cross diffusion: only w.r.t. species of the same compartment
    for (int w : M.comps[S.comp].species) {
      std::string dp = p + ".cross_diffusion." + M.spec[w].name;
      if (!M.ini.has_sub(dp)) continue;
      Diff D;
      D.wrt = w;
      if (!read_diff(M, dp, D.tensor, D.d)) continue;
      for (int k : M.comps[S.comp].species) {
        std::string jp = dp + ".jacobian." + M.spec[k].name;
        if (!M.ini.has_sub(jp)) continue;
        Diff::DJac J;
        J.wrt = k;
        if (read_diff(M, jp, J.tensor, J.d)) D.jac.push_back(J);
      }
      S.diff.push_back(D);
    }
    for (size_t c = 0; c < M.comps.size(); ++c) {
      std::string op = p + ".outflow." + M.comps[c].name;
      if (!M.ini.has_sub(op)) continue;
      Outflow O;
      O.to_comp = (int)c;
      O.e = cexpr(M, op + ".expression");
      if (O.e.null()) continue;
      O.jac = jac_list(M, op);
      S.outflow.push_back(O);
      M.has_outflow = true;
    }
  }
}

static void mark_compartments(Model& M) {
  Mesh& m = M.mesh;
  M.elem_comp.assign(m.nelem, -1);
  std::vector<Expr> ce;
  for (auto& C : M.comps) {
    std::string e = M.ini.get("compartments." + C.name + ".expression", "");
    std::string type = M.ini.get("compartments." + C.name + ".type", "expression");
    if (type != "expression") throw std::runtime_error("oracle: compartment type '" + type + "' not supported");
    ce.push_back(compile(e, M.scope));
  }
  std::vector<double> v(V_SPEC + 4 * M.spec.size() + 4, 0.0);
  for (long e = 0; e < m.nelem; ++e) {
    const long* c = &m.conn[e * m.nvpe];
    double cen[3] = {0, 0, 0};
    for (int a = 0; a < m.nvpe; ++a)
      for (int k = 0; k < 3; ++k) cen[k] += m.X[3 * c[a] + k] / m.nvpe;
    for (int k = 0; k < 3; ++k) v[V_POS + k] = cen[k];
    if (!m.phys.empty()) v[V_SPEC + 4 * M.spec.size()] = m.phys[e];
    int found = -1;
    for (size_t i = 0; i < ce.size(); ++i) {
      if (!ce[i].null() && ce[i](v.data()) != 0.0) {
        if (found >= 0)
          throw std::runtime_error("oracle: overlapping compartments are not supported");
        found = (int)i;
      }
    }
    M.elem_comp[e] = found;
  }
  for (long e = 0; e < m.nelem; ++e)
    if (M.elem_comp[e] >= 0) M.comps[M.elem_comp[e]].elems.push_back(e);
}

static void build_layout(Model& M) {
  Mesh& m = M.mesh;
  long dof = 0, slot = 0;
  for (auto& C : M.comps) {
    C.g2l.assign(m.nvert, -1);
    std::vector<char> used(m.nvert, 0);
    if (C.b > 0)
      for (long e : C.elems)
        for (int a = 0; a < m.nvpe; ++a) used[m.conn[e * m.nvpe + a]] = 1;
    for (long v = 0; v < m.nvert; ++v)
      if (used[v]) {
        C.g2l[v] = (long)C.nodes.size();
        C.nodes.push_back(v);
      }
    C.dof_off = dof;
    C.slot_off = slot;
    dof += (long)C.nodes.size() * C.b;
    slot += (long)C.nodes.size();
  }
  M.ndof = dof;
  M.nslot = slot;
  M.slot_dof.resize(slot);
  M.slot_b.resize(slot);
  for (auto& C : M.comps)
    for (size_t n = 0; n < C.nodes.size(); ++n) {
      M.slot_dof[C.slot_off + n] = C.dof_off + (long)n * C.b;
      M.slot_b[C.slot_off + n] = C.b;
    }
}

static void build_facets(Model& M) {
  Mesh& m = M.mesh;
  auto F = faces(m);
  std::map<std::vector<long>, std::pair<long, int>> open;
  std::vector<std::tuple<long, int, long>> pairs;  // e1, f1, e2 (-1 boundary)
  for (long e = 0; e < m.nelem; ++e) {
    for (size_t f = 0; f < F.size(); ++f) {
      std::vector<long> key;
      for (int l : F[f]) key.push_back(m.conn[e * m.nvpe + l]);
      std::sort(key.begin(), key.end());
      auto it = open.find(key);
      if (it == open.end()) {
        open[key] = {e, (int)f};
      } else {
        pairs.emplace_back(it->second.first, it->second.second, e);
        open.erase(it);
      }
    }
  }
  for (auto& kv : open) pairs.emplace_back(kv.second.first, kv.second.second, -1);
  for (auto& p : pairs) {
    long e1 = std::get<0>(p), e2 = std::get<2>(p);
    int c1 = M.elem_comp[e1];
    int c2 = e2 >= 0 ? M.elem_comp[e2] : -1;
    if (e2 < 0) {
      if (c1 >= 0) M.facets.push_back({e1, -1, std::get<1>(p), c1, c1});
    } else if (c1 != c2 && c1 >= 0 && c2 >= 0) {
      M.facets.push_back({e1, e2, std::get<1>(p), c1, c2});
    }
  }
}

static void build_pattern(Model& M) {
  Mesh& m = M.mesh;
  std::vector<std::vector<long>> nb(M.nslot);
  for (auto& C : M.comps) {
    if (C.b == 0) continue;
    for (long e : C.elems) {
      const long* c = &m.conn[e * m.nvpe];
      for (int a = 0; a < m.nvpe; ++a)
        for (int b = 0; b < m.nvpe; ++b) nb[C.slot_off + C.g2l[c[a]]].push_back(C.slot_off + C.g2l[c[b]]);
    }
  }
  // interface couplings between the two sides of a facet
  auto F = faces(m);
  for (auto& f : M.facets) {
    if (f.e_out < 0) continue;
    auto& C1 = M.comps[f.c_in];
    auto& C2 = M.comps[f.c_out];
    if (C1.b == 0 || C2.b == 0) continue;
    for (int la : F[f.f_in]) {
      long va = m.conn[f.e_in * m.nvpe + la];
      for (int lb : F[f.f_in]) {
        long vb = m.conn[f.e_in * m.nvpe + lb];
        long s1 = C1.slot_off + C1.g2l[va], s2 = C2.slot_off + C2.g2l[vb];
        nb[s1].push_back(s2);
        nb[s2].push_back(s1);
      }
    }
  }
  M.adj_ptr.assign(M.nslot + 1, 0);
  for (long s = 0; s < M.nslot; ++s) {
    auto& l = nb[s];
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
    M.adj_ptr[s + 1] = M.adj_ptr[s] + (long)l.size();
  }
  M.adj.resize(M.adj_ptr[M.nslot]);
  M.adj_off.resize(M.adj_ptr[M.nslot]);
  for (long s = 0; s < M.nslot; ++s) {
    long off = 0;
    for (size_t k = 0; k < nb[s].size(); ++k) {
      M.adj[M.adj_ptr[s] + k] = nb[s][k];
      M.adj_off[M.adj_ptr[s] + k] = off;
      off += M.slot_b[nb[s][k]];
    }
    std::vector<long>().swap(nb[s]);
  }
  // scalar CSR rows: species rows of a slot share its neighbour list
  M.row_ptr.assign(M.ndof + 1, 0);
  for (long s = 0; s < M.nslot; ++s) {
    long len = 0;
    for (long k = M.adj_ptr[s]; k < M.adj_ptr[s + 1]; ++k) len += M.slot_b[M.adj[k]];
    for (int i = 0; i < M.slot_b[s]; ++i) M.row_ptr[M.slot_dof[s] + i + 1] = len;
  }
  for (long r = 0; r < M.ndof; ++r) M.row_ptr[r + 1] += M.row_ptr[r];
  M.col.resize(M.row_ptr[M.ndof]);
  for (long s = 0; s < M.nslot; ++s)
    for (int i = 0; i < M.slot_b[s]; ++i) {
      long p = M.row_ptr[M.slot_dof[s] + i];
      for (long k = M.adj_ptr[s]; k < M.adj_ptr[s + 1]; ++k)
        for (int j = 0; j < M.slot_b[M.adj[k]]; ++j) M.col[p++] = (int)(M.slot_dof[M.adj[k]] + j);
    }
}

// greedy vertex-disjoint element coloring per compartment (the reference's
// colored concurrent assembly, diffusion_reaction/make_step_operator.hh:66-121)
static void build_colors(Model& M) {
  Mesh& m = M.mesh;
  std::vector<uint64_t> used(m.nvert, 0);
  M.colors.clear();
  for (auto& C : M.comps) {
    if (C.b == 0) continue;
    std::vector<std::vector<long>> cl;
    for (long e : C.elems) {
      uint64_t mask = 0;
      for (int a = 0; a < m.nvpe; ++a) mask |= used[m.conn[e * m.nvpe + a]];
      int c = 0;
      while (c < 64 && ((mask >> c) & 1)) ++c;
      if (c == 64) throw std::runtime_error("oracle: coloring overflow");
      if ((int)cl.size() <= c) cl.resize(c + 1);
      cl[c].push_back(e);
      for (int a = 0; a < m.nvpe; ++a) used[m.conn[e * m.nvpe + a]] |= (1ull << c);
    }
    for (auto& l : cl) M.colors.push_back(l);
    std::fill(used.begin(), used.end(), 0);
  }
}

static long find_adj(const Model& M, long s, long t) {
  auto b = M.adj.begin() + M.adj_ptr[s], e = M.adj.begin() + M.adj_ptr[s + 1];
  auto it = std::lower_bound(b, e, t);
  if (it == e || *it != t) return -1;
  return it - M.adj.begin();
}

// ---------------------------------------------------------------- assembly
// Local element kernel.  Evaluates, for one element of compartment C,
//   rl[a*b+i]                   += cm * r^M + ca * r^A
//   Jl[((a*b+i)*nv + c)*b + j]  += cm * dr^M + ca * dr^A
// with the weak form of local_operator.hh:479-485 (residual) and :605-700
// (Jacobian), test index (a,i), trial index (c,j).
struct Work {
  std::vector<double> v;
  std::vector<double> xl, rl, Jl, rp, xp;
};

static void volume_local(const Model& M, const Compartment& C, long e, const double* xl, double t, double cm,
                         double ca, double* rl, double* Jl, Work& w) {
  const Mesh& m = M.mesh;
  int d = m.dim, nv = m.nvpe, b = C.b;
  Geo g = element_geo(m, e);
  QRule q = m.cube ? cube_rule(d) : simplex_rule(d);
  double vol = 0;
  for (double wq : q.w) vol += wq * g.det;
  std::fill(w.v.begin(), w.v.end(), 0.0);
  double* v = w.v.data();
  v[V_TIME] = t;
  v[V_EVOL] = vol;
  v[V_INVOL] = 1.0;
  double phi[8], dref[8][3], grad[8][3];
  for (size_t iq = 0; iq < q.w.size(); ++iq) {
    shape(m, q.x[iq].data(), phi, dref);
    for (int a = 0; a < nv; ++a)
      for (int k = 0; k < 3; ++k) {
        double s = 0;
        for (int l = 0; l < d; ++l) s += dref[a][l] * g.Jinv[l][k];
        grad[a][k] = k < d ? s : 0.0;
      }
    double fac = q.w[iq] * g.det;
    for (int k = 0; k < 3; ++k) {
      double s = g.x0[k];
      for (int l = 0; l < d; ++l) s += g.J[k][l] * q.x[iq][l];
      v[V_POS + k] = k < d ? s : 0.0;
    }
    for (int i = 0; i < b; ++i) {
      int gs = C.species[i];
      double val = 0, gr[3] = {0, 0, 0};
      for (int a = 0; a < nv; ++a) {
        val += xl[a * b + i] * phi[a];
        for (int k = 0; k < 3; ++k) gr[k] += xl[a * b + i] * grad[a][k];
      }
      v[V_SPEC + 4 * gs] = val;
      for (int k = 0; k < 3; ++k) v[V_SPEC + 4 * gs + 1 + k] = gr[k];
    }
    for (int i = 0; i < b; ++i) {
      const Species& S = M.spec[C.species[i]];
      double ui = v[V_SPEC + 4 * C.species[i]];
      double scal = 0, flux[3] = {0, 0, 0};
      // mass form: storage * u
      double stg = 0;
      if (cm != 0.0 && !S.storage.null()) {
        stg = S.storage(v);
        scal += cm * stg * ui;
      }
      double rea = 0;
      if (ca != 0.0 && !S.reaction.null()) {
        rea = S.reaction(v);
        scal -= ca * rea;
      }
      double vel[3] = {0, 0, 0};
      bool hasv = false;
      if (ca != 0.0)
        for (int k = 0; k < d; ++k)
          if (!S.velocity[k].null()) {
            vel[k] = S.velocity[k](v);
            hasv = true;
          }
      if (hasv)
        for (int k = 0; k < d; ++k) flux[k] += ca * vel[k] * ui;
      // cross diffusion: flux -= D_ij grad u_j
      std::vector<std::array<std::array<double, 3>, 3>> Dv(S.diff.size());
      if (ca != 0.0)
        for (size_t n = 0; n < S.diff.size(); ++n) {
          const Diff& D = S.diff[n];
          auto& T = Dv[n];
          for (auto& r : T) r = {0, 0, 0};
          if (D.tensor) {
            for (int k = 0; k < d; ++k)
              for (int l = 0; l < d; ++l) T[k][l] = D.d[k][l].null() ? 0.0 : D.d[k][l](v);
          } else {
            double s = D.d[0][0](v);
            for (int k = 0; k < d; ++k) T[k][k] = s;
          }
          const double* gj = &v[V_SPEC + 4 * D.wrt + 1];
          for (int k = 0; k < d; ++k) {
            double s = 0;
            for (int l = 0; l < d; ++l) s += T[k][l] * gj[l];
            flux[k] -= ca * s;
          }
        }
      for (int a = 0; a < nv; ++a) {
        double dd = 0;
        for (int k = 0; k < d; ++k) dd += flux[k] * grad[a][k];
        rl[a * b + i] += (scal * phi[a] - dd) * fac;
      }
      if (!Jl) continue;
      auto J = [&](int a, int c, int j) -> double& { return Jl[((a * b + i) * nv + c) * b + j]; };
      // storage: S phi_c psi_a (j == i) + dS/du_k u_i phi_c psi_a
      if (cm != 0.0 && !S.storage.null()) {
        for (int a = 0; a < nv; ++a)
          for (int c = 0; c < nv; ++c) J(a, c, i) += cm * stg * phi[c] * phi[a] * fac;
        for (auto& jj : S.storage_jac) {
          if (M.spec[jj.wrt].comp != S.comp) continue;
          double dv = jj.e(v) * ui;
          int j = M.spec[jj.wrt].local;
          for (int a = 0; a < nv; ++a)
            for (int c = 0; c < nv; ++c) J(a, c, j) += cm * dv * phi[c] * phi[a] * fac;
        }
      }
      if (ca == 0.0) continue;
      if (!S.reaction.null())
        for (auto& jj : S.reaction_jac) {
          if (M.spec[jj.wrt].comp != S.comp) continue;
          double dr = jj.e(v);
          int j = M.spec[jj.wrt].local;
          for (int a = 0; a < nv; ++a)
            for (int c = 0; c < nv; ++c) J(a, c, j) -= ca * dr * phi[c] * phi[a] * fac;
        }
      if (hasv) {
        // d/du_c of -(v u_i).grad psi_a = -(v phi_c).grad psi_a  (exact)
        for (int a = 0; a < nv; ++a)
          for (int c = 0; c < nv; ++c) {
            double s = 0;
            for (int k = 0; k < d; ++k) s += vel[k] * grad[a][k];
            J(a, c, i) -= ca * s * phi[c] * fac;
          }
        for (size_t n = 0; n < S.vel_jac.size(); ++n) {
          int w2 = S.vel_jac_wrt[n];
          if (M.spec[w2].comp != S.comp) continue;
          int j = M.spec[w2].local;
          double dvk[3] = {0, 0, 0};
          for (int k = 0; k < d; ++k) dvk[k] = S.vel_jac[n][k].null() ? 0.0 : S.vel_jac[n][k](v);
          for (int a = 0; a < nv; ++a) {
            double s = 0;
            for (int k = 0; k < d; ++k) s += dvk[k] * ui * grad[a][k];
            for (int c = 0; c < nv; ++c) J(a, c, j) -= ca * s * phi[c] * fac;
          }
        }
      }
      for (size_t n = 0; n < S.diff.size(); ++n) {
        const Diff& D = S.diff[n];
        auto& T = Dv[n];
        int j = M.spec[D.wrt].local;
        for (int a = 0; a < nv; ++a)
          for (int c = 0; c < nv; ++c) {
            double s = 0;
            for (int k = 0; k < d; ++k)
              for (int l = 0; l < d; ++l) s += T[k][l] * grad[c][l] * grad[a][k];
            J(a, c, j) += ca * s * fac;
          }
        // nonlinear diffusion: (dD_ij/du_k grad u_j) phi_c . grad psi_a  (exact)
        for (auto& DJ : D.jac) {
          int k2 = M.spec[DJ.wrt].local;
          double T2[3][3] = {{0}};
          if (DJ.tensor) {
            for (int k = 0; k < d; ++k)
              for (int l = 0; l < d; ++l) T2[k][l] = DJ.d[k][l].null() ? 0.0 : DJ.d[k][l](v);
          } else {
            double s = DJ.d[0][0](v);
            for (int k = 0; k < d; ++k) T2[k][k] = s;
          }
          const double* gj = &v[V_SPEC + 4 * D.wrt + 1];
          double fl[3] = {0, 0, 0};
          for (int k = 0; k < d; ++k)
            for (int l = 0; l < d; ++l) fl[k] += T2[k][l] * gj[l];
          for (int a = 0; a < nv; ++a) {
            double s = 0;
            for (int k = 0; k < d; ++k) s += fl[k] * grad[a][k];
            for (int c = 0; c < nv; ++c) J(a, c, k2) += ca * s * phi[c] * fac;
          }
        }
      }
    }
  }
}

// facet geometry: quadrature points on the facet given by local face vertices
struct FacetQ {
  std::vector<std::array<double, 3>> x;  // global points
  std::vector<double> fac;               // weight * integration element
  double normal[3];
  double volume;
};

static FacetQ facet_quad(const Mesh& m, long e, const std::vector<int>& fl) {
  FacetQ r;
  int d = m.dim;
  const long* c = &m.conn[e * m.nvpe];
  std::vector<std::array<double, 3>> P;
  for (int l : fl) P.push_back({m.X[3 * c[l]], m.X[3 * c[l] + 1], m.X[3 * c[l] + 2]});
  // spanning vectors: simplices P1-P0, P2-P0; cube faces (lexicographic) P1-P0, P2-P0
  std::array<double, 3> t1{0, 0, 0}, t2{0, 0, 0};
  if (d >= 2)
    for (int k = 0; k < 3; ++k) t1[k] = P[1][k] - P[0][k];
  if (d == 3)
    for (int k = 0; k < 3; ++k) t2[k] = P[2][k] - P[0][k];
  double ie;
  double n[3] = {0, 0, 0};
  if (d == 1) {
    ie = 1.0;
    n[0] = 1.0;
  } else if (d == 2) {
    ie = std::sqrt(t1[0] * t1[0] + t1[1] * t1[1]);
    n[0] = t1[1] / ie;
    n[1] = -t1[0] / ie;
  } else {
    double cr[3] = {t1[1] * t2[2] - t1[2] * t2[1], t1[2] * t2[0] - t1[0] * t2[2], t1[0] * t2[1] - t1[1] * t2[0]};
    ie = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    for (int k = 0; k < 3; ++k) n[k] = cr[k] / ie;
  }
  // orient away from the element's centroid
  double ec[3] = {0, 0, 0}, fcn[3] = {0, 0, 0};
  for (int a = 0; a < m.nvpe; ++a)
    for (int k = 0; k < 3; ++k) ec[k] += m.X[3 * c[a] + k] / m.nvpe;
  for (auto& p : P)
    for (int k = 0; k < 3; ++k) fcn[k] += p[k] / P.size();
  double s = 0;
  for (int k = 0; k < 3; ++k) s += n[k] * (fcn[k] - ec[k]);
  if (s < 0)
    for (int k = 0; k < 3; ++k) n[k] = -n[k];
  for (int k = 0; k < 3; ++k) r.normal[k] = n[k];
  QRule q = m.cube ? cube_rule(d - 1) : simplex_rule(d - 1);
  double vol = 0;
  for (size_t i = 0; i < q.w.size(); ++i) {
    std::array<double, 3> x;
    for (int k = 0; k < 3; ++k) x[k] = P[0][k] + q.x[i][0] * t1[k] + q.x[i][1] * t2[k];
    r.x.push_back(x);
    r.fac.push_back(q.w[i] * ie);
    vol += q.w[i] * ie;
  }
  r.volume = vol;
  return r;
}

// reference coordinates of a global point in an affine element
static void local_coords(const Mesh& m, const Geo& g, const double* x, double* xi) {
  for (int l = 0; l < m.dim; ++l) {
    double s = 0;
    for (int k = 0; k < m.dim; ++k) s += g.Jinv[l][k] * (x[k] - g.x0[k]);
    xi[l] = s;
  }
  for (int l = m.dim; l < 3; ++l) xi[l] = 0;
}

// Transmission / boundary outflow on one facet (stiffness form only,
// local_operator.hh:777-971).  Species of the inside compartment take values
// from the inside element, species of the outside compartment from the outside
// element; each side's outflow is tested with its own side's basis.  Writes
// residual into rin/rout and the four Jacobian blocks (exact derivatives).
struct FacetOut {
  std::vector<double> rin, rout;      // nv*b_in, nv*b_out
  std::vector<double> Jii, Jio, Joi, Joo;
};

static void facet_local(const Model& M, const Facet& F, const std::vector<std::vector<int>>& FL, const double* xin,
                        const double* xout, double t, double ca, FacetOut& o, std::vector<double>& vbuf) {
  const Mesh& m = M.mesh;
  int nv = m.nvpe, d = m.dim;
  const Compartment& Ci = M.comps[F.c_in];
  const Compartment& Co = M.comps[F.c_out];
  bool bnd = F.e_out < 0;
  int bi = Ci.b, bo = Co.b;
  o.rin.assign(nv * bi, 0.0);
  o.rout.assign(nv * bo, 0.0);
  o.Jii.assign(nv * bi * nv * bi, 0.0);
  o.Jio.assign(nv * bi * nv * bo, 0.0);
  o.Joi.assign(nv * bo * nv * bi, 0.0);
  o.Joo.assign(nv * bo * nv * bo, 0.0);
  FacetQ fq = facet_quad(m, F.e_in, FL[F.f_in]);
  Geo gi = element_geo(m, F.e_in);
  Geo go = bnd ? gi : element_geo(m, F.e_out);
  long eo = bnd ? F.e_in : F.e_out;
  double* v = vbuf.data();
  std::fill(vbuf.begin(), vbuf.end(), 0.0);
  v[V_TIME] = t;
  v[V_EVOL] = fq.volume;
  v[V_INBND] = bnd ? 1.0 : 0.0;
  v[V_INSKEL] = bnd ? 0.0 : 1.0;
  double phi_i[8], phi_o[8], dref[8][3], gri[8][3], gro[8][3];
  for (size_t q = 0; q < fq.x.size(); ++q) {
    const double* x = fq.x[q].data();
    double xi[3], xo[3];
    local_coords(m, gi, x, xi);
    local_coords(m, go, x, xo);
    shape(m, xi, phi_i, dref);
    for (int a = 0; a < nv; ++a)
      for (int k = 0; k < 3; ++k) {
        double s = 0;
        for (int l = 0; l < d; ++l) s += dref[a][l] * gi.Jinv[l][k];
        gri[a][k] = k < d ? s : 0;
      }
    shape(m, xo, phi_o, dref);
    for (int a = 0; a < nv; ++a)
      for (int k = 0; k < 3; ++k) {
        double s = 0;
        for (int l = 0; l < d; ++l) s += dref[a][l] * go.Jinv[l][k];
        gro[a][k] = k < d ? s : 0;
      }
    for (int k = 0; k < 3; ++k) v[V_POS + k] = k < d ? x[k] : 0.0;
    // values of both sides
    auto fill = [&](const Compartment& C, const double* xl, const double* phi, double (*gr)[3]) {
      for (int i = 0; i < C.b; ++i) {
        int gs = C.species[i];
        double val = 0, g3[3] = {0, 0, 0};
        for (int a = 0; a < nv; ++a) {
          val += xl[a * C.b + i] * phi[a];
          for (int k = 0; k < 3; ++k) g3[k] += xl[a * C.b + i] * gr[a][k];
        }
        v[V_SPEC + 4 * gs] = val;
        for (int k = 0; k < 3; ++k) v[V_SPEC + 4 * gs + 1 + k] = g3[k];
      }
    };
    fill(Ci, xin, phi_i, gri);
    if (!bnd) fill(Co, xout, phi_o, gro);
    double fac = fq.fac[q];
    // inside species: outflow towards the outside compartment (or own on boundary)
    auto side = [&](const Compartment& Cs, int target, const double* phi_s, const double* nrm, std::vector<double>& r,
                    std::vector<double>& Jss, std::vector<double>& Jsx, const Compartment& Cx, const double* phi_x,
                    bool self_is_in) {
      for (int k = 0; k < 3; ++k) v[V_NORMAL + k] = k < d ? nrm[k] : 0.0;
      for (int i = 0; i < Cs.b; ++i) {
        const Species& S = M.spec[Cs.species[i]];
        for (auto& O : S.outflow) {
          if (O.to_comp != target) continue;
          double val = O.e(v);
          for (int a = 0; a < nv; ++a) r[a * Cs.b + i] += ca * val * phi_s[a] * fac;
          for (auto& jj : O.jac) {
            int wc = M.spec[jj.wrt].comp;
            double dv = jj.e(v);
            int j = M.spec[jj.wrt].local;
            bool same = (&M.comps[wc] == &Cs);
            bool other = (!bnd && &M.comps[wc] == &Cx);
            if (same) {
              for (int a = 0; a < nv; ++a)
                for (int c = 0; c < nv; ++c) Jss[((a * Cs.b + i) * nv + c) * Cs.b + j] += ca * dv * phi_s[c] * phi_s[a] * fac;
            } else if (other) {
              for (int a = 0; a < nv; ++a)
                for (int c = 0; c < nv; ++c) Jsx[((a * Cs.b + i) * nv + c) * Cx.b + j] += ca * dv * phi_x[c] * phi_s[a] * fac;
            }
          }
        }
      }
      (void)self_is_in;
    };
    if (bnd) {
      side(Ci, F.c_in, phi_i, fq.normal, o.rin, o.Jii, o.Jio, Co, phi_o, true);
    } else {
      side(Ci, F.c_out, phi_i, fq.normal, o.rin, o.Jii, o.Jio, Co, phi_o, true);
      double nn[3] = {-fq.normal[0], -fq.normal[1], -fq.normal[2]};
      side(Co, F.c_in, phi_o, nn, o.rout, o.Joo, o.Joi, Ci, phi_i, false);
    }
  }
  (void)eo;
}

// Global assembly: r = cm*R^M(x) + ca*R^A(x) (+= when accumulate), and, when
// vals != nullptr, the Jacobian values on the scalar CSR pattern.  Elements are
// processed colour by colour; inside a colour no two elements share a vertex,
// so threads never write the same entry.
static void assemble(Model& M, const double* x, double t, double cm, double ca, double* r, double* vals,
                     bool accumulate) {
  Mesh& m = M.mesh;
  int nv = m.nvpe;
  if (!accumulate) std::fill(r, r + M.ndof, 0.0);
  if (vals) std::fill(vals, vals + M.row_ptr[M.ndof], 0.0);
  size_t nvar = V_SPEC + 4 * M.spec.size() + 4;
  // element -> compartment map for the colour lists
  for (auto& col : M.colors) {
    long n = (long)col.size();
#pragma omp parallel
    {
      Work w;
      w.v.assign(nvar, 0.0);
#pragma omp for schedule(static)
      for (long idx = 0; idx < n; ++idx) {
        long e = col[idx];
        const Compartment& C = M.comps[M.elem_comp[e]];
        int b = C.b;
        int nl = nv * b;
        w.xl.assign(nl, 0.0);
        w.rl.assign(nl, 0.0);
        const long* cn = &m.conn[e * nv];
        long sl[8];
        for (int a = 0; a < nv; ++a) {
          sl[a] = C.slot_off + C.g2l[cn[a]];
          for (int i = 0; i < b; ++i) w.xl[a * b + i] = x[M.slot_dof[sl[a]] + i];
        }
        double* Jl = nullptr;
        if (vals) {
          w.Jl.assign((size_t)nl * nl, 0.0);
          Jl = w.Jl.data();
        }
        if (vals && M.numerical_jac && !M.is_linear) {
          // forward differences of the element residual, local_operator.hh:713-765
          volume_local(M, C, e, w.xl.data(), t, cm, ca, w.rl.data(), nullptr, w);
          w.xp = w.xl;
          for (int c = 0; c < nv; ++c)
            for (int j = 0; j < b; ++j) {
              w.rp.assign(nl, 0.0);
              double delta = M.fd_eps * (1.0 + std::fabs(w.xl[c * b + j]));
              w.xp[c * b + j] += delta;
              volume_local(M, C, e, w.xp.data(), t, cm, ca, w.rp.data(), nullptr, w);
              w.xp[c * b + j] = w.xl[c * b + j];
              for (int a = 0; a < nv; ++a)
                for (int i = 0; i < b; ++i)
                  Jl[((a * b + i) * nv + c) * b + j] = (w.rp[a * b + i] - w.rl[a * b + i]) / delta;
            }
        } else {
          volume_local(M, C, e, w.xl.data(), t, cm, ca, w.rl.data(), Jl, w);
        }
        for (int a = 0; a < nv; ++a) {
          long s = sl[a];
          for (int i = 0; i < b; ++i) r[M.slot_dof[s] + i] += w.rl[a * b + i];
          if (!vals) continue;
          for (int c = 0; c < nv; ++c) {
            long k = find_adj(M, s, sl[c]);
            long off = M.adj_off[k];
            for (int i = 0; i < b; ++i) {
              long p = M.row_ptr[M.slot_dof[s] + i] + off;
              for (int j = 0; j < b; ++j) vals[p + j] += Jl[((a * b + i) * nv + c) * b + j];
            }
          }
        }
      }
    }
  }
  if (!M.has_outflow || ca == 0.0) return;
  // facets: sequential (few of them)
  auto FL = faces(m);
  std::vector<double> vbuf(nvar, 0.0);
  FacetOut o;
  for (auto& F : M.facets) {
    const Compartment& Ci = M.comps[F.c_in];
    const Compartment& Co = M.comps[F.c_out];
    bool bnd = F.e_out < 0;
    if (Ci.b == 0 && (bnd || Co.b == 0)) continue;
    std::vector<double> xin(nv * Ci.b), xout(nv * Co.b, 0.0);
    long si[8], so[8];
    for (int a = 0; a < nv; ++a) {
      long vi = m.conn[F.e_in * nv + a];
      si[a] = Ci.b ? Ci.slot_off + Ci.g2l[vi] : -1;
      for (int i = 0; i < Ci.b; ++i) xin[a * Ci.b + i] = x[M.slot_dof[si[a]] + i];
      if (!bnd) {
        long vo = m.conn[F.e_out * nv + a];
        so[a] = Co.b ? Co.slot_off + Co.g2l[vo] : -1;
        for (int i = 0; i < Co.b; ++i) xout[a * Co.b + i] = x[M.slot_dof[so[a]] + i];
      }
    }
    auto run = [&](const double* xi_, const double* xo_, FacetOut& out) {
      facet_local(M, F, FL, xi_, xo_, t, ca, out, vbuf);
    };
    if (vals && M.numerical_jac && !M.is_linear) {
      // numerical skeleton Jacobian: forward differences on both sides
      run(xin.data(), xout.data(), o);
      FacetOut base = o, up;
      auto fdcol = [&](std::vector<double>& xs, int idx, bool side_in) {
        double keep = xs[idx];
        double delta = M.fd_eps * (1.0 + std::fabs(keep));
        xs[idx] += delta;
        run(xin.data(), xout.data(), up);
        xs[idx] = keep;
        int c = idx / (side_in ? std::max(1, Ci.b) : std::max(1, Co.b));
        int j = idx % (side_in ? std::max(1, Ci.b) : std::max(1, Co.b));
        for (int a = 0; a < nv; ++a) {
          for (int i = 0; i < Ci.b; ++i) {
            double dv = (up.rin[a * Ci.b + i] - base.rin[a * Ci.b + i]) / delta;
            if (side_in) o.Jii[((a * Ci.b + i) * nv + c) * Ci.b + j] = dv;
            else o.Jio[((a * Ci.b + i) * nv + c) * Co.b + j] = dv;
          }
          for (int i = 0; i < Co.b && !bnd; ++i) {
            double dv = (up.rout[a * Co.b + i] - base.rout[a * Co.b + i]) / delta;
            if (side_in) o.Joi[((a * Co.b + i) * nv + c) * Ci.b + j] = dv;
            else o.Joo[((a * Co.b + i) * nv + c) * Co.b + j] = dv;
          }
        }
      };
      for (int idx = 0; idx < nv * Ci.b; ++idx) fdcol(xin, idx, true);
      if (!bnd)
        for (int idx = 0; idx < nv * Co.b; ++idx) fdcol(xout, idx, false);
      o.rin = base.rin;
      o.rout = base.rout;
    } else {
      run(xin.data(), xout.data(), o);
    }
    auto scatter = [&](const long* rs, int br, const long* cs, int bc, const std::vector<double>& Jb) {
      for (int a = 0; a < nv; ++a)
        for (int c = 0; c < nv; ++c) {
          long k = find_adj(M, rs[a], cs[c]);
          if (k < 0) continue;  // vertex off the facet: zero block by construction
          long off = M.adj_off[k];
          for (int i = 0; i < br; ++i) {
            long p = M.row_ptr[M.slot_dof[rs[a]] + i] + off;
            for (int j = 0; j < bc; ++j) vals[p + j] += Jb[((a * br + i) * nv + c) * bc + j];
          }
        }
    };
    for (int a = 0; a < nv; ++a)
      for (int i = 0; i < Ci.b; ++i) r[M.slot_dof[si[a]] + i] += o.rin[a * Ci.b + i];
    if (!bnd)
      for (int a = 0; a < nv; ++a)
        for (int i = 0; i < Co.b; ++i) r[M.slot_dof[so[a]] + i] += o.rout[a * Co.b + i];
    if (!vals) continue;
    if (Ci.b) scatter(si, Ci.b, si, Ci.b, o.Jii);
    if (!bnd && Ci.b && Co.b) {
      scatter(si, Ci.b, so, Co.b, o.Jio);
      scatter(so, Co.b, si, Ci.b, o.Joi);
    }
    if (!bnd && Co.b) scatter(so, Co.b, so, Co.b, o.Joo);
  }
}

// ---------------------------------------------------------------- linear algebra
struct CSR {
  long n;
  const long* rp;
  const int* ci;
  double* v;
};

// The Krylov solve is sequential, as istl's BiCGSTABSolver is; with
// krylov_threads its matvecs, dots and vector updates run on all threads
// (a stronger CPU baseline; dots then sum in another order).
static bool krylov_threads = false;
static bool par(long n) { return krylov_threads && n > 20000; }

static void spmv(const CSR& A, const double* x, double* y) {
#pragma omp parallel for schedule(static) if (par(A.n))
  for (long i = 0; i < A.n; ++i) {
    double s = 0;
    for (long p = A.rp[i]; p < A.rp[i + 1]; ++p) s += A.v[p] * x[A.ci[p]];
    y[i] = s;
  }
}

static double dot(long n, const double* a, const double* b) {
  double s = 0;
#pragma omp parallel for schedule(static) reduction(+ : s) if (par(n))
  for (long i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// istl SeqILU with n = 0: in-place incomplete LU on the matrix pattern
// (IKJ variant), unit lower factor, the upper factor keeps the diagonal.
struct ILU0 {
  CSR A;
  std::vector<double> f;
  std::vector<long> diag;
  void factor(const CSR& src) {
    A = src;
    f.assign(src.v, src.v + src.rp[src.n]);
    diag.assign(A.n, -1);
    for (long i = 0; i < A.n; ++i)
      for (long p = A.rp[i]; p < A.rp[i + 1]; ++p)
        if (A.ci[p] == i) diag[i] = p;
    for (long i = 0; i < A.n; ++i) {
      if (diag[i] < 0) throw std::runtime_error("oracle: ILU0 needs a diagonal entry");
      for (long p = A.rp[i]; p < A.rp[i + 1] && A.ci[p] < i; ++p) {
        long k = A.ci[p];
        f[p] /= f[diag[k]];
        double lik = f[p];
        // a_ij -= l_ik * u_kj for j > k present in both rows
        long q = diag[k] + 1, pe = A.rp[i + 1];
        long r = p + 1;
        for (; q < A.rp[k + 1]; ++q) {
          int j = A.ci[q];
          while (r < pe && A.ci[r] < j) ++r;
          if (r < pe && A.ci[r] == j) f[r] -= lik * f[q];
        }
      }
    }
  }
  void apply(const double* b, double* x) const {
    for (long i = 0; i < A.n; ++i) {
      double s = b[i];
      for (long p = A.rp[i]; p < diag[i]; ++p) s -= f[p] * x[A.ci[p]];
      x[i] = s;
    }
    for (long i = A.n - 1; i >= 0; --i) {
      double s = x[i];
      for (long p = diag[i] + 1; p < A.rp[i + 1]; ++p) s -= f[p] * x[A.ci[p]];
      x[i] = s / f[diag[i]];
    }
  }
};

struct Precond {
  std::string type;
  ILU0 ilu;
  std::vector<double> dinv;
  void setup(const CSR& A, const std::string& t) {
    type = t;
    if (t == "ILU" || t == "BlockJacobi") {
      ilu.factor(A);
    } else if (t == "Jacobi") {
      dinv.assign(A.n, 0.0);
      for (long i = 0; i < A.n; ++i)
        for (long p = A.rp[i]; p < A.rp[i + 1]; ++p)
          if (A.ci[p] == i) dinv[i] = 1.0 / A.v[p];
    } else if (t != "None" && t != "Richardson") {
      // Richardson: istl's SeqRichardson with its default weight 1, the
      // identity (the preconditioner of the matrix-free solve)
      throw std::runtime_error("oracle: preconditioner '" + t + "' not supported");
    }
  }
  void apply(long n, const double* b, double* x) const {
    if (type == "ILU" || type == "BlockJacobi") ilu.apply(b, x);
    else if (type == "Jacobi") {
#pragma omp parallel for schedule(static) if (par(n))
      for (long i = 0; i < n; ++i) x[i] = dinv[i] * b[i];
    } else {
      std::copy(b, b + n, x);
    }
  }
};

// Preconditioned BiCGSTAB (van der Vorst), solving A x = b from x = 0; stops
// when ||r|| <= rtol * ||r0||, checking after both half steps as istl does.
static int bicgstab(const CSR& A, const Precond& P, const double* b, double* x, double rtol, int maxit, int* its,
                    double* red) {
  long n = A.n;
  std::vector<double> r(b, b + n), rh(b, b + n), p(n, 0.0), v(n, 0.0), ph(n), s(n), sh(n), t(n);
  std::fill(x, x + n, 0.0);
  double n0 = std::sqrt(dot(n, r.data(), r.data()));
  *its = 0;
  *red = 0;
  if (n0 == 0.0) return 0;
  double rho = 1, alpha = 1, omega = 1;
  for (int it = 1; it <= maxit; ++it) {
    double rho1 = dot(n, rh.data(), r.data());
    if (rho1 == 0.0) return 4;
    double beta = (rho1 / rho) * (alpha / omega);
#pragma omp parallel for schedule(static) if (par(n))
    for (long i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    P.apply(n, p.data(), ph.data());
    spmv(A, ph.data(), v.data());
    double rv = dot(n, rh.data(), v.data());
    if (rv == 0.0) return 4;
    alpha = rho1 / rv;
#pragma omp parallel for schedule(static) if (par(n))
    for (long i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
    double ns = std::sqrt(dot(n, s.data(), s.data()));
    if (ns <= rtol * n0) {
      for (long i = 0; i < n; ++i) x[i] += alpha * ph[i];
      *its = it;
      *red = ns / n0;
      return 0;
    }
    P.apply(n, s.data(), sh.data());
    spmv(A, sh.data(), t.data());
    double tt = dot(n, t.data(), t.data());
    omega = tt == 0.0 ? 0.0 : dot(n, t.data(), s.data()) / tt;
#pragma omp parallel for schedule(static) if (par(n))
    for (long i = 0; i < n; ++i) {
      x[i] += alpha * ph[i] + omega * sh[i];
      r[i] = s[i] - omega * t[i];
    }
    rho = rho1;
    double nr = std::sqrt(dot(n, r.data(), r.data()));
    *its = it;
    *red = nr / n0;
    if (!std::isfinite(nr)) return 3;
    if (nr <= rtol * n0) return 0;
    if (omega == 0.0) return 4;
  }
  return 2;
}

// ---------------------------------------------------------------- time step
static double nl_norm(const Model& M, const std::vector<double>& r) {
  if (M.nl_norm == "l_inf") {
    double m = 0;
    for (double v : r) m = std::max(m, std::fabs(v));
    return m;
  }
  if (M.nl_norm == "l_1") {
    double s = 0;
    for (double v : r) s += std::fabs(v);
    return s;
  }
  return dot((long)r.size(), r.data(), r.data());  // two_norm2, make_step_operator.hh:274-276
}

struct StepStats {
  int newton = 0, linear = 0;
};

// One Runge-Kutta step of size dt from (M.u, M.time).  Stage r solves
//   sum_{j<r} [a_rj M(u_j) + dt b_rj A(u_j, t + d_j dt)] + a_rr M(x) + dt b_rr A(x, t + d_r dt) = 0
// with Newton (or one defect correction when the model is linear).
static int step(Model& M, double dt, StepStats& st) {
  long n = M.ndof;
  const RK& K = M.rk;
  std::vector<std::vector<double>> U(K.stages + 1);
  U[0] = M.u;
  std::vector<double> hist(n), r(n), z(n), vals(M.row_ptr[n]);
  CSR A{n, M.row_ptr.data(), M.col.data(), vals.data()};
  auto t0 = std::chrono::steady_clock::now();
  double tasm = 0, tsol = 0;
  for (int s = 1; s <= K.stages; ++s) {
    std::fill(hist.begin(), hist.end(), 0.0);
    for (int j = 0; j < s; ++j) {
      double am = K.a[s - 1][j], ab = dt * K.b[s - 1][j];
      if (am == 0.0 && ab == 0.0) continue;
      assemble(M, U[j].data(), M.time + K.d[j] * dt, am, ab, hist.data(), nullptr, true);
    }
    double cm = K.a[s - 1][s], ca = dt * K.b[s - 1][s];
    double ts = M.time + K.d[s] * dt;
    std::vector<double> x = U[s - 1];
    auto F = [&](bool jac) {
      auto a0 = std::chrono::steady_clock::now();
      r = hist;
      assemble(M, x.data(), ts, cm, ca, r.data(), jac ? vals.data() : nullptr, true);
      tasm += std::chrono::duration<double>(std::chrono::steady_clock::now() - a0).count();
    };
    auto solve = [&](double rtol) -> int {
      auto a0 = std::chrono::steady_clock::now();
      Precond P;
      P.setup(A, M.precond);  // rebuilt every call, make_step_operator.hh:120-127
      int its;
      double red;
      int rc = bicgstab(A, P, r.data(), z.data(), rtol, M.lin_maxit, &its, &red);
      st.linear += its;
      tsol += std::chrono::duration<double>(std::chrono::steady_clock::now() - a0).count();
      return rc;
    };
    if (M.is_linear) {
      F(true);
      int rc = solve(M.lin_rtol);
      if (rc) return rc == 2 ? 2 : rc;
      for (long i = 0; i < n; ++i) x[i] -= z[i];
      st.newton += 1;
    } else {
      // classic PDELab Newton control: stop below max(abs, rel * d0); linear
      // reduction max(stop/(10 d_k), min(min_red, d_k^2/d_{k-1}^2)) unless
      // fixed; re-linearize unless the defect rate stays below the threshold
      F(false);
      double nrm0 = nl_norm(M, r), nrm = nrm0, prev = nrm0;
      const double stop = std::max(M.nl_atol, M.nl_rtol * nrm0);
      int it = 0;
      bool have_jac = false;
      for (;;) {
        if (!std::isfinite(nrm)) return 3;
        bool conv = nrm <= stop || nrm == 0.0;
        if (conv && it >= M.nl_min) break;
        if (it >= M.nl_max) return 1;
        bool relin = !have_jac || M.lin_threshold <= 0.0 || nrm / prev > M.lin_threshold;
        if (relin) {
          std::vector<double> keep = r;
          F(true);  // assembles the residual again with the Jacobian
          have_jac = true;
          (void)keep;
        }
        double rtol = M.lin_rtol;
        if (!M.fixed_lin_tol) {
          double ratio = it == 0 ? 1.0 : (nrm * nrm) / (prev * prev);
          rtol = std::max(stop / (10.0 * nrm), std::min(M.min_lin_red, ratio));
        }
        int rc = solve(rtol);
        if (rc == 3 || rc == 4) return rc;
        if (rc == 2) return 2;
        for (long i = 0; i < n; ++i) x[i] -= z[i];
        ++it;
        F(false);
        prev = nrm;
        nrm = nl_norm(M, r);
      }
      st.newton += it;
    }
    U[s] = x;
  }
  (void)t0;
  M.last_assembly_s = tasm;
  M.last_solve_s = tsol;
  M.u = U[K.stages];
  M.time += dt;
  return 0;
}

// ---------------------------------------------------------------- time loop
// SimpleAdaptiveStepper::do_step (stepper.hh:330-366) inside TimeStepper::
// evolve and snap_to_time (stepper.hh:145-240): halve (decrease factor) and
// retry on failure, grow by the increase factor (clamped to the maximum)
// after success, stop 2 dt before the end and snap with equal steps.  Times
// are compared with Dune::FloatCmp's relative-weak default (8 machine
// epsilons of the larger magnitude).
struct EvolveRecord {
  std::vector<double> accepted;
  int failed = 0;
};

static bool t_eq(double a, double b) {
  return std::fabs(a - b) <= 8.0 * std::numeric_limits<double>::epsilon() * std::max(std::fabs(a), std::fabs(b));
}

static int evolve(Model& M, double t_end, double& dt, double dec, double inc, double dmin, double dmax, bool snap,
                  EvolveRecord& rec) {
  auto out_of_range = [&](double h) {
    if (dmin > 0.0 && std::fabs(h) < std::fabs(dmin) && !t_eq(std::fabs(h), std::fabs(dmin))) return true;
    if (dmax > 0.0 && std::fabs(h) > std::fabs(dmax) && !t_eq(std::fabs(h), std::fabs(dmax))) return true;
    return false;
  };
  auto adaptive = [&](double& h) -> int {
    if (out_of_range(h)) return 5;
    StepStats st;
    int rc = step(M, h, st);
    while (rc) {
      rec.failed++;
      h = h * dec;
      if (out_of_range(h)) return 5;
      StepStats st2;
      rc = step(M, h, st2);
    }
    rec.accepted.push_back(h);
    double next = h * inc;
    if (dmax > 0.0) next = std::min(std::max(next, -std::fabs(dmax)), std::fabs(dmax));
    h = next;
    return 0;
  };
  const double stop = snap ? 2.0 : 1.0;
  for (;;) {
    const double a = M.time + stop * dt;
    if (!(a < t_end || t_eq(a, t_end))) break;
    int rc = adaptive(dt);
    if (rc) return rc;
  }
  if (!snap) return 0;
  int tries = 0, rc = 0;
  while (M.time < t_end && !t_eq(M.time, t_end)) {
    const int n = (int)std::ceil((t_end - M.time) / dt);
    if (n <= 0) throw std::runtime_error("Timestep doesn't make advances towards snap step");
    dt = (t_end - M.time) / n;
    rc = adaptive(dt);
    if (rc) {
      if (tries++ == 100) throw std::runtime_error("Snapping time exceeded maximum iteration count");
      dt *= 0.5;
    }
  }
  return rc;
}


// ---------------------------------------------------------------- transform / reduce
// reduce.hh:40-285 of the reference, as the CLI runs it (sequential policy,
// model_single_compartment.impl.hh:247): per key, the evaluation at every
// order-4 point of every element in grid order, folded from initial.value by
// reduction (default +); that result is reduced once more with initial.value
// (reduce.hh:212-217); then transformation, error and warn.  Only position,
// time, integration_factor, the species and their gradients are bound
// (reduce.hh:171-190); the other point keywords keep their zero defaults
// (local_domain.hh:40-46).  Elements outside every compartment are skipped.

// Order-4 rules: Gauss-Legendre, 3 points per axis, on boxes; on triangles the
// 6-point rule with area coordinates (a, a, 1 - 2a); on tetrahedra the
// 11-point rule (centroid, orbit of 4, orbit of 6).
static QRule reduce_rule(int d, bool cube) {
  QRule r;
  if (cube || d == 1) {
    const double s = 0.5 * std::sqrt(0.6);
    const double x1[3] = {0.5 - s, 0.5, 0.5 + s}, w1[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};
    int n = 1;
    for (int k = 0; k < d; ++k) n *= 3;
    for (int i = 0; i < n; ++i) {
      std::array<double, 3> p{0, 0, 0};
      double w = 1.0;
      int m = i;
      for (int k = 0; k < d; ++k) {
        p[k] = x1[m % 3];
        w *= w1[m % 3];
        m /= 3;
      }
      r.x.push_back(p);
      r.w.push_back(w);
    }
    return r;
  }
  if (d == 2) {
    const double a[2] = {0.445948490915965, 0.091576213509771};
    const double w[2] = {0.5 * 0.223381589678011, 0.5 * 0.109951743655322};
    for (int o = 0; o < 2; ++o) {
      const double b = 1.0 - 2.0 * a[o];
      r.x.push_back({a[o], a[o], 0});
      r.x.push_back({b, a[o], 0});
      r.x.push_back({a[o], b, 0});
      for (int k = 0; k < 3; ++k) r.w.push_back(w[o]);
    }
    return r;
  }
  r.x.push_back({0.25, 0.25, 0.25});
  r.w.push_back(-74.0 / 5625.0);
  const double a = 1.0 / 14.0, a3 = 11.0 / 14.0;
  const double pa[4][3] = {{a, a, a}, {a3, a, a}, {a, a3, a}, {a, a, a3}};
  for (auto& q : pa) {
    r.x.push_back({q[0], q[1], q[2]});
    r.w.push_back(343.0 / 45000.0);
  }
  const double b = 0.25 * (1.0 + std::sqrt(5.0 / 14.0)), c = 0.25 * (1.0 - std::sqrt(5.0 / 14.0));
  const double pb[6][3] = {{b, b, c}, {b, c, b}, {c, b, b}, {c, c, b}, {c, b, c}, {b, c, c}};
  for (auto& q : pb) {
    r.x.push_back({q[0], q[1], q[2]});
    r.w.push_back(56.0 / 2250.0);
  }
  return r;
}

// "a, b: body" with the arguments in slots 0.. of their own value array
struct RFun {
  bool present = false;
  int nargs = 0;
  Expr body;
};

static RFun reduce_function(const std::string& text, const Scope& model_scope, int nargs) {
  RFun f;
  f.present = true;
  auto colon = text.find(':');
  if (colon == std::string::npos) throw std::runtime_error("oracle: expected 'args: expression' in '" + text + "'");
  Scope sc;
  sc.consts = model_scope.consts;
  sc.funcs = model_scope.funcs;
  std::string args = text.substr(0, colon);
  std::vector<std::string> names;
  std::string cur;
  for (char ch : args + ",") {
    if (ch == ',') {
      std::string t = trim(cur);
      if (!t.empty()) names.push_back(t);
      cur.clear();
    } else {
      cur += ch;
    }
  }
  if ((int)names.size() != nargs) throw std::runtime_error("oracle: function needs " + std::to_string(nargs) + " arguments");
  for (int i = 0; i < nargs; ++i) {
    sc.consts.erase(names[i]);
    sc.vars[names[i]] = i;
  }
  f.nargs = nargs;
  f.body = compile(text.substr(colon + 1), sc);
  return f;
}

static double call(const RFun& f, double a, double b = 0.0) {
  const double v[2] = {a, b};
  return f.body.null() ? 0.0 : f.body(v);
}

struct ReduceOut {
  std::vector<std::string> names;
  std::vector<double> values;
  std::vector<int> flags;  // 0, 1 warn, 2 error
};

static int reduce(const Model& M, const Ini& cfg, const std::string& prefix, ReduceOut& out) {
  struct Key {
    std::string name;
    Expr eval;
    RFun red, tr, err, warn;
    double init = 0;
  };
  std::vector<Key> keys;
  const std::string p0 = prefix.empty() ? "" : prefix + ".";
  for (auto& k : cfg.subkeys(prefix)) {
    Key K;
    K.name = k;
    const std::string p = p0 + k + ".";
    if (cfg.has(p + "evaluation.expression")) K.eval = compile(cfg.get(p + "evaluation.expression"), M.scope);
    if (cfg.has(p + "reduction.expression")) K.red = reduce_function(cfg.get(p + "reduction.expression"), M.scope, 2);
    K.init = cfg.get_double(p + "initial.value", 0.0);
    if (cfg.has(p + "transformation.expression"))
      K.tr = reduce_function(cfg.get(p + "transformation.expression"), M.scope, 1);
    if (cfg.has(p + "error.expression")) K.err = reduce_function(cfg.get(p + "error.expression"), M.scope, 1);
    if (cfg.has(p + "warn.expression")) K.warn = reduce_function(cfg.get(p + "warn.expression"), M.scope, 1);
    keys.push_back(std::move(K));
  }
  const Mesh& m = M.mesh;
  const int d = m.dim, nv = m.nvpe;
  const QRule q = reduce_rule(d, m.cube);
  std::vector<double> acc(keys.size());
  for (size_t k = 0; k < keys.size(); ++k) acc[k] = keys[k].init;
  std::vector<double> v(V_SPEC + 4 * M.spec.size() + 4, 0.0);
  double phi[8], dref[8][3], grad[8][3];
  for (long e = 0; e < m.nelem; ++e) {
    const int c = M.elem_comp[e];
    if (c < 0) continue;
    const Compartment& C = M.comps[c];
    Geo g = element_geo(m, e);
    std::fill(v.begin(), v.end(), 0.0);
    v[V_TIME] = M.time;
    const long* conn = &m.conn[e * nv];
    for (size_t iq = 0; iq < q.w.size(); ++iq) {
      shape(m, q.x[iq].data(), phi, dref);
      for (int a = 0; a < nv; ++a)
        for (int k = 0; k < 3; ++k) {
          double s = 0;
          for (int l = 0; l < d; ++l) s += dref[a][l] * g.Jinv[l][k];
          grad[a][k] = k < d ? s : 0.0;
        }
      for (int k = 0; k < 3; ++k) {
        double s = g.x0[k];
        for (int l = 0; l < d; ++l) s += g.J[k][l] * q.x[iq][l];
        v[V_POS + k] = k < d ? s : 0.0;
      }
      v[V_IFAC] = q.w[iq] * g.det;
      for (int i = 0; i < C.b; ++i) {
        const int gs = C.species[i];
        double val = 0, gr[3] = {0, 0, 0};
        for (int a = 0; a < nv; ++a) {
          const long node = C.g2l[conn[a]];
          const double x = M.u[C.dof_off + node * C.b + i];
          val += x * phi[a];
          for (int k = 0; k < 3; ++k) gr[k] += x * grad[a][k];
        }
        v[V_SPEC + 4 * gs] = val;
        for (int k = 0; k < 3; ++k) v[V_SPEC + 4 * gs + 1 + k] = gr[k];
      }
      for (size_t k = 0; k < keys.size(); ++k) {
        const double ev = keys[k].eval.null() ? 0.0 : keys[k].eval(v.data());
        acc[k] = keys[k].red.present ? call(keys[k].red, ev, acc[k]) : ev + acc[k];
      }
    }
  }
  int rc = 0;
  for (size_t k = 0; k < keys.size(); ++k) {
    const Key& K = keys[k];
    double val = K.red.present ? call(K.red, acc[k], K.init) : acc[k] + K.init;
    if (K.tr.present) val = call(K.tr, val);
    int flag = 0;
    if (K.err.present && call(K.err, val) != 0.0) {
      flag = 2;
      rc = 1;
    } else if (K.warn.present && call(K.warn, val) != 0.0) {
      flag = 1;
    }
    out.names.push_back(K.name);
    out.values.push_back(val);
    out.flags.push_back(flag);
  }
  return rc;
}

static void configure(Model& M, const std::string& ini, const std::string& over) {
  M.ini.parse(ini);
  if (!over.empty()) M.ini.parse_overrides(over);
  const Ini& I = M.ini;
  Mesh& m = M.mesh;
  m.dim = (int)I.get_double("grid.dimension", 2);
  if (m.dim < 1 || m.dim > 3) throw std::runtime_error("oracle: grid.dimension must be 1, 2 or 3");
  auto cells = I.get_doubles("grid.cells");
  auto ext = I.get_doubles("grid.extensions");
  auto org = I.get_doubles("grid.origin");
  for (int k = 0; k < m.dim; ++k) {
    m.nc[k] = k < (int)cells.size() ? (long)cells[k] : 1;
    m.ext[k] = k < (int)ext.size() ? ext[k] : 1.0;
    m.origin[k] = k < (int)org.size() ? org[k] : 0.0;
  }
  int lvl = (int)I.get_double("grid.refinement_level", 0);
  for (int k = 0; k < m.dim; ++k) m.nc[k] <<= lvl;
  std::string ct = I.get("grid.cell_type", "simplex");
  if (ct != "simplex" && ct != "cube") throw std::runtime_error("oracle: grid.cell_type must be simplex or cube");
  m.cube = (ct == "cube");
  if (I.has("grid.path")) {
    read_gmsh(m, I.get("grid.path"));
    refine_file_mesh(m, lvl);
  } else {
    build_mesh(m);
  }
  build_species(M);
  build_scope(M);
  if (!m.phys.empty()) M.scope.vars["gmsh_id"] = V_SPEC + 4 * (int)M.spec.size();
  build_equations(M);
  mark_compartments(M);
  build_layout(M);
  if (M.has_outflow) build_facets(M);
  build_pattern(M);
  build_colors(M);
  M.u.assign(M.ndof, 0.0);
  std::string p = "model.time_step_operator.";
  M.time = I.get_double(p + "time_begin", 0.0);
  M.rk = rk_table(I.get(p + "type", "Alexander2"));
  std::string jt = I.get("model.jacobian.type", "analytical");
  if (jt != "analytical" && jt != "numerical")
    throw std::runtime_error("The option 'model.jacobian.type' must be either 'analytical' or 'numerical'");
  M.numerical_jac = (jt == "numerical");
  M.fd_eps = I.get_double("model.jacobian.epsilon", 1e-7);
  std::string il = I.get("model.is_linear", "false");
  M.is_linear = (il == "true" || il == "1");
  M.lin_type = I.get(p + "linear_solver.type", "BiCGSTAB");
  if (M.lin_type != "BiCGSTAB") throw std::runtime_error("oracle: only BiCGSTAB is restated");
  std::string pt = I.get(p + "linear_solver.preconditioner.type", "ILU");
  if (pt == "BlockJacobi") {
    std::string inner = I.get(p + "linear_solver.preconditioner.diagonal_solver.preconditioner.type", "ILU");
    pt = (inner == "ILU") ? "ILU" : inner;
  }
  M.precond = pt;
  M.lin_rtol = I.get_double(p + "linear_solver.convergence_condition.relative_tolerance", 1e-4);
  auto lr = I.get_doubles(p + "linear_solver.convergence_condition.iteration_range");
  if (lr.size() == 2) M.lin_maxit = (int)lr[1];
  M.nl_rtol = I.get_double(p + "nonlinear_solver.convergence_condition.relative_tolerance", 1e-4);
  M.nl_atol = I.get_double(p + "nonlinear_solver.convergence_condition.absolute_tolerance", 0.0);
  auto nr = I.get_doubles(p + "nonlinear_solver.convergence_condition.iteration_range");
  if (nr.size() == 2) {
    M.nl_min = (int)nr[0];
    M.nl_max = (int)nr[1];
  }
  M.nl_norm = I.get(p + "nonlinear_solver.norm", "l_2");
  std::string fx = I.get(p + "nonlinear_solver.dx_inverse_fixed_tolerance", "false");
  M.fixed_lin_tol = (fx == "true" || fx == "1");
  M.min_lin_red = I.get_double(p + "nonlinear_solver.dx_inverse_min_relative_tolerance", 0.1);
  M.lin_threshold = I.get_double(p + "nonlinear_solver.linearization_threshold", 0.0);
}

}  // namespace orc

// ---------------------------------------------------------------- C ABI
using orc::Model;

static thread_local std::string g_err;

static int fail(const std::exception& e) {
  g_err = e.what();
  return -1;
}

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void orc_set_threads(int n) {
  orc::nthreads_req = n;
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#endif
}

void orc_set_krylov_threads(int on) { orc::krylov_threads = on != 0; }

// transform/reduce on the current state: config NULL or "" reads the model's
// model.reduce keys; names are written '\n'-separated.  Returns 0, 1 when an
// error expression fired, -1 on failure.
int orc_reduce(void* h, const char* config, char* names, int names_len, double* values, int* flags, int capacity,
               int* count) {
  try {
    auto* M = static_cast<Model*>(h);
    orc::ReduceOut out;
    int rc;
    if (config && *config) {
      orc::Ini cfg;
      cfg.parse(config);
      rc = orc::reduce(*M, cfg, "", out);
    } else {
      rc = orc::reduce(*M, M->ini, "model.reduce", out);
    }
    std::string joined;
    for (size_t i = 0; i < out.names.size(); ++i) {
      if (i) joined += "\n";
      joined += out.names[i];
      if ((int)i < capacity) {
        values[i] = out.values[i];
        flags[i] = out.flags[i];
      }
    }
    std::snprintf(names, names_len, "%s", joined.c_str());
    *count = (int)out.names.size();
    return rc;
  } catch (const std::exception& e) {
    fail(e);
    return -1;
  }
}

int orc_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void* orc_create(const char* ini, const char* overrides) {
  try {
    auto* M = new Model();
    orc::configure(*M, ini ? ini : "", overrides ? overrides : "");
    return M;
  } catch (const std::exception& e) {
    fail(e);
    return nullptr;
  }
}

void orc_destroy(void* h) { delete static_cast<Model*>(h); }

long orc_n_dof(void* h) { return static_cast<Model*>(h)->ndof; }
int orc_n_species(void* h) { return (int)static_cast<Model*>(h)->spec.size(); }
int orc_n_compartments(void* h) { return (int)static_cast<Model*>(h)->comps.size(); }
long orc_n_vertices(void* h) { return static_cast<Model*>(h)->mesh.nvert; }
long orc_n_elements(void* h) { return static_cast<Model*>(h)->mesh.nelem; }
long orc_n_facets(void* h) { return (long)static_cast<Model*>(h)->facets.size(); }

int orc_species(void* h, int s, char* name, int len, int* comp, int* local) {
  auto* M = static_cast<Model*>(h);
  if (s < 0 || s >= (int)M->spec.size()) return -1;
  std::snprintf(name, len, "%s", M->spec[s].name.c_str());
  *comp = M->spec[s].comp;
  *local = M->spec[s].local;
  return 0;
}

// number of nodes of a compartment; fills their global vertex ids if ids != 0
long orc_comp_nodes(void* h, int c, long* ids) {
  auto* M = static_cast<Model*>(h);
  auto& C = M->comps[c];
  if (ids) std::copy(C.nodes.begin(), C.nodes.end(), ids);
  return (long)C.nodes.size();
}
long orc_comp_dof_offset(void* h, int c) { return static_cast<Model*>(h)->comps[c].dof_off; }
int orc_comp_species_count(void* h, int c) { return static_cast<Model*>(h)->comps[c].b; }

int orc_vertex_coords(void* h, double* xyz) {
  auto* M = static_cast<Model*>(h);
  std::copy(M->mesh.X.begin(), M->mesh.X.end(), xyz);
  return 0;
}

int orc_set_state(void* h, const double* u, double t) {
  auto* M = static_cast<Model*>(h);
  std::copy(u, u + M->ndof, M->u.begin());
  M->time = t;
  return 0;
}

int orc_get_state(void* h, double* u, double* t) {
  auto* M = static_cast<Model*>(h);
  std::copy(M->u.begin(), M->u.end(), u);
  if (t) *t = M->time;
  return 0;
}

// One time step; returns 0 on success, 1 Newton not converged, 2 Krylov not
// converged, 3 non-finite, 4 breakdown, -1 error.  On failure the state is
// unchanged.
int orc_step(void* h, double dt, int* newton_its, int* linear_its) {
  auto* M = static_cast<Model*>(h);
  try {
    orc::StepStats st;
    int rc = orc::step(*M, dt, st);
    if (newton_its) *newton_its = st.newton;
    if (linear_its) *linear_its = st.linear;
    return rc;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

int orc_evolve(void* h, double t_end, double dt, double dec, double inc, double dmin, double dmax, int snap,
               double* accepted, int capacity, int* n_accepted, int* n_failed, double* dt_next) {
  auto* M = static_cast<Model*>(h);
  try {
    orc::EvolveRecord rec;
    int rc = orc::evolve(*M, t_end, dt, dec, inc, dmin, dmax, snap != 0, rec);
    if (n_accepted) *n_accepted = (int)rec.accepted.size();
    if (n_failed) *n_failed = rec.failed;
    if (dt_next) *dt_next = dt;
    for (int k = 0; k < (int)rec.accepted.size() && k < capacity; ++k) accepted[k] = rec.accepted[k];
    return rc;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

double orc_last_assembly_seconds(void* h) { return static_cast<Model*>(h)->last_assembly_s; }
double orc_last_solve_seconds(void* h) { return static_cast<Model*>(h)->last_solve_s; }

long orc_nnz(void* h) {
  auto* M = static_cast<Model*>(h);
  return M->row_ptr[M->ndof];
}

int orc_pattern(void* h, long* rowptr, int* col) {
  auto* M = static_cast<Model*>(h);
  std::copy(M->row_ptr.begin(), M->row_ptr.end(), rowptr);
  std::copy(M->col.begin(), M->col.end(), col);
  return 0;
}

// r = cm*R^M(x,t) + ca*R^A(x,t); vals (nullable) = its Jacobian on orc_pattern
int orc_assemble(void* h, const double* x, double t, double cm, double ca, double* r, double* vals) {
  auto* M = static_cast<Model*>(h);
  try {
    orc::assemble(*M, x, t, cm, ca, r, vals, false);
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

// Solve A x = b with the model's preconditioner and BiCGSTAB.
int orc_solve(void* h, const double* vals, const double* b, double* x, double rtol, int maxit, int* its) {
  auto* M = static_cast<Model*>(h);
  try {
    std::vector<double> v(vals, vals + M->row_ptr[M->ndof]);
    orc::CSR A{M->ndof, M->row_ptr.data(), M->col.data(), v.data()};
    orc::Precond P;
    P.setup(A, M->precond);
    double red;
    return orc::bicgstab(A, P, b, x, rtol, maxit, its, &red);
  } catch (const std::exception& e) {
    return fail(e);
  }
}

// Per-species integral of x over its compartment (global species order); the
// element rules are exact for P1/Q1 integrands of degree <= 1 per axis.
int orc_mass(void* h, const double* x, double* out) {
  auto* M = static_cast<Model*>(h);
  auto& m = M->mesh;
  std::fill(out, out + M->spec.size(), 0.0);
  orc::QRule q = m.cube ? orc::cube_rule(m.dim) : orc::simplex_rule(m.dim);
  double phi[8], dref[8][3];
  for (auto& C : M->comps) {
    if (!C.b) continue;
    for (long e : C.elems) {
      orc::Geo g = orc::element_geo(m, e);
      for (size_t iq = 0; iq < q.w.size(); ++iq) {
        orc::shape(m, q.x[iq].data(), phi, dref);
        for (int a = 0; a < m.nvpe; ++a) {
          long s = C.slot_off + C.g2l[m.conn[e * m.nvpe + a]];
          for (int i = 0; i < C.b; ++i) out[C.species[i]] += q.w[iq] * g.det * phi[a] * x[M->slot_dof[s] + i];
        }
      }
    }
  }
  return 0;
}

}  // extern "C"

// Evaluate one expression with named variables (parser tests).  `context` is
// INI text whose [parser_context] constants/functions are visible.
extern "C" int orc_eval(const char* context, const char* text, const char* const* names, const double* values, int n,
                        double* out) {
  try {
    orc::Model M;
    M.ini.parse(context ? context : "");
    orc::Scope sc;
    for (int i = 0; i < n; ++i) sc.vars[names[i]] = i;
    // constants and functions of the context, as build_scope reads them
    for (auto& k : M.ini.subkeys("parser_context")) {
      std::string p = "parser_context." + k;
      std::string type = M.ini.get(p + ".type", "");
      if (type == "constant") sc.consts[k] = std::stod(M.ini.get(p + ".value"));
      if (type == "function") {
        std::string e = M.ini.get(p + ".expression");
        auto colon = e.find(':');
        orc::UserFunction f;
        std::string args = e.substr(0, colon);
        size_t st = 0;
        while (st <= args.size()) {
          size_t c = args.find(',', st);
          std::string a = orc::trim(args.substr(st, c == std::string::npos ? std::string::npos : c - st));
          if (!a.empty()) f.args.push_back(a);
          if (c == std::string::npos) break;
          st = c + 1;
        }
        f.body = e.substr(colon + 1);
        sc.funcs[k] = f;
      }
      if (type == "interpolation") sc.tables[k] = orc::read_table(M.ini, p);
    }
    orc::Expr e = orc::compile(text, sc);
    *out = e.null() ? 0.0 : e(values);
    return e.null() ? 1 : 0;
  } catch (const std::exception& ex) {
    g_err = ex.what();
    return -1;
  }
}
