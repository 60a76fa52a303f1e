#include "layout.hh"

#include <algorithm>
#include <map>
#include <fstream>
#include <cmath>
#include <numeric>
#include <unordered_map>

#include "error.hh"

namespace dcopasi {

// ------------------------------------------------------------------ grid
Grid make_grid(int dim, const std::vector<double>& cells, const std::vector<double>& ext,
               const std::vector<double>& origin, int refine, bool cube) {
  if (dim < 1 || dim > 3) throw IOError("grid.dimension must be 1, 2 or 3");
  Grid g;
  g.dim = dim;
  g.cube = cube || dim == 1;
  for (int k = 0; k < 3; ++k) {
    g.n[k] = 1;
    g.h[k] = 1.0;
    g.origin[k] = 0.0;
  }
  for (int k = 0; k < dim; ++k) {
    long c = k < static_cast<int>(cells.size()) ? static_cast<long>(cells[k]) : 1;
    if (c < 1) throw IOError("grid.cells must be positive");
    g.n[k] = c << refine;
    double e = k < static_cast<int>(ext.size()) ? ext[k] : 1.0;
    g.h[k] = e / static_cast<double>(g.n[k]);
    g.origin[k] = k < static_cast<int>(origin.size()) ? origin[k] : 0.0;
  }
  g.nvert = 1;
  g.ncell = 1;
  for (int k = 0; k < dim; ++k) {
    g.nvert *= g.n[k] + 1;
    g.ncell *= g.n[k];
  }
  if (g.nvert >= (1L << 31) / 2) throw NotImplemented("grid too large for 32-bit slot indices");
  g.ncorner = 1 << dim;
  if (g.cube) {
    g.nclass = 1;
    g.nvpe = g.ncorner;
    std::array<int, 8> c{};
    for (int a = 0; a < 8; ++a) c[a] = a;
    g.corner.push_back(c);
  } else {
    // Kuhn split: one simplex per axis permutation, path corner 0 -> 2^d-1
    std::vector<int> p(dim);
    std::iota(p.begin(), p.end(), 0);
    g.nvpe = dim + 1;
    do {
      std::array<int, 8> c{};
      int bits = 0;
      c[0] = 0;
      for (int s = 0; s < dim; ++s) {
        bits |= 1 << p[s];
        c[s + 1] = bits;
      }
      g.corner.push_back(c);
    } while (std::next_permutation(p.begin(), p.end()));
    g.nclass = static_cast<int>(g.corner.size());
  }
  return g;
}

Grid read_gmsh(const std::string& path, int dim) {
  if (dim < 1 || dim > 3) throw IOError("grid.dimension must be 1, 2 or 3");
  std::ifstream in(path);
  if (!in) throw IOError("cannot open grid file '" + path + "'");
  std::string tok;
  std::unordered_map<long, std::array<double, 3>> nodes;
  struct Elem {
    std::array<long, 4> v;
    double phys;
  };
  std::vector<Elem> elems;
  const int cell_type = dim == 1 ? 1 : dim == 2 ? 2 : 4;  // line, triangle, tetrahedron
  const int nvpe = dim + 1;
  bool have_format = false;
  while (in >> tok) {
    if (tok == "$MeshFormat") {
      double version;
      int file_type, data_size;
      in >> version >> file_type >> data_size;
      if (version < 2.0 || version >= 3.0) throw NotImplemented("gmsh format " + std::to_string(version) + " (2.x only)");
      if (file_type != 0) throw NotImplemented("binary gmsh files");
      have_format = true;
    } else if (tok == "$Nodes") {
      long n;
      in >> n;
      for (long i = 0; i < n; ++i) {
        long id;
        std::array<double, 3> x;
        in >> id >> x[0] >> x[1] >> x[2];
        nodes[id] = x;
      }
    } else if (tok == "$Elements") {
      long n;
      in >> n;
      for (long i = 0; i < n; ++i) {
        long id;
        int type, ntags;
        in >> id >> type >> ntags;
        std::vector<long> tags(ntags);
        for (auto& t : tags) in >> t;
        static const int nv_of[16] = {0, 2, 3, 4, 4, 8, 6, 5, 3, 6, 9, 10, 27, 18, 14, 1};
        if (type < 1 || type > 15) throw NotImplemented("gmsh element type " + std::to_string(type));
        std::vector<long> v(nv_of[type]);
        for (auto& x : v) in >> x;
        const int edim = (type == 15) ? 0 : (type == 1 || type == 8) ? 1 : (type == 2 || type == 3 || type == 9 || type == 10) ? 2 : 3;
        if (edim != dim) continue;  // lower-dimensional elements are boundary segments
        if (type != cell_type) throw NotImplemented("gmsh cells of type " + std::to_string(type) + " (simplices only)");
        Elem e;
        e.v = {-1, -1, -1, -1};
        for (int a = 0; a < nvpe; ++a) e.v[a] = v[a];
        e.phys = ntags > 0 ? static_cast<double>(tags[0]) : 0.0;
        elems.push_back(e);
      }
    }
    if (!in && !in.eof()) throw IOError("malformed gmsh file '" + path + "'");
  }
  if (!have_format) throw IOError("'" + path + "' has no $MeshFormat section");
  if (elems.empty()) throw IOError("'" + path + "' has no cells of dimension " + std::to_string(dim));
  // used nodes in increasing node number
  std::vector<long> used;
  for (const auto& e : elems)
    for (int a = 0; a < nvpe; ++a) used.push_back(e.v[a]);
  std::sort(used.begin(), used.end());
  used.erase(std::unique(used.begin(), used.end()), used.end());
  std::unordered_map<long, int> index;
  Grid g;
  g.dim = dim;
  g.cube = false;
  g.unstructured = true;
  g.nvpe = g.ncorner = nvpe;
  g.nclass = 1;
  std::array<int, 8> c{};
  for (int a = 0; a < 8; ++a) c[a] = a;
  g.corner.push_back(c);
  g.nvert = static_cast<long>(used.size());
  g.ncell = static_cast<long>(elems.size());
  if (g.nvert >= (1L << 31) / 2) throw NotImplemented("grid too large for 32-bit slot indices");
  g.X.resize(3 * used.size());
  for (size_t i = 0; i < used.size(); ++i) {
    auto it = nodes.find(used[i]);
    if (it == nodes.end()) throw IOError("gmsh cell uses undefined node " + std::to_string(used[i]));
    index[used[i]] = static_cast<int>(i);
    for (int k = 0; k < 3; ++k) g.X[3 * i + k] = k < dim ? it->second[k] : 0.0;
  }
  for (const auto& e : elems) {
    for (int a = 0; a < nvpe; ++a) g.conn.push_back(index[e.v[a]]);
    g.gmsh_id.push_back(e.phys);
  }
  return g;
}

void refine_simplices(Grid& g, int levels) {
  if (!g.unstructured) throw Error("refine_simplices: grid is not unstructured");
  const int nv = g.nvpe;
  // local edges in creation order
  static const int E1[1][2] = {{0, 1}};
  static const int E2[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  static const int E3[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  const int ne = g.dim == 1 ? 1 : g.dim == 2 ? 3 : 6;
  const int(*E)[2] = g.dim == 1 ? E1 : g.dim == 2 ? E2 : E3;
  for (int level = 0; level < levels; ++level) {
    std::map<std::pair<int, int>, int> mid;
    std::vector<int> conn;
    std::vector<double> gid;
    conn.reserve(g.conn.size() * (g.dim == 3 ? 8 : g.dim == 2 ? 4 : 2));
    for (long c = 0; c < g.ncell; ++c) {
      const int* v = &g.conn[static_cast<size_t>(c) * nv];
      int m[6];
      for (int e = 0; e < ne; ++e) {
        const int a = v[E[e][0]], b = v[E[e][1]];
        const auto key = std::make_pair(std::min(a, b), std::max(a, b));
        auto it = mid.find(key);
        if (it == mid.end()) {
          const int id = static_cast<int>(g.X.size() / 3);
          for (int k = 0; k < 3; ++k) g.X.push_back(0.5 * (g.X[3 * a + k] + g.X[3 * b + k]));
          it = mid.emplace(key, id).first;
        }
        m[e] = it->second;
      }
      auto child = [&](std::initializer_list<int> vs) {
        for (int x : vs) conn.push_back(x);
        gid.push_back(g.gmsh_id[c]);
      };
      if (g.dim == 1) {
        child({v[0], m[0]});
        child({m[0], v[1]});
      } else if (g.dim == 2) {
        const int m01 = m[0], m02 = m[1], m12 = m[2];
        child({v[0], m01, m02});
        child({m01, v[1], m12});
        child({m02, m12, v[2]});
        child({m01, m12, m02});
      } else {
        const int m01 = m[0], m02 = m[1], m03 = m[2], m12 = m[3], m13 = m[4], m23 = m[5];
        child({v[0], m01, m02, m03});
        child({m01, v[1], m12, m13});
        child({m02, m12, v[2], m23});
        child({m03, m13, m23, v[3]});
        auto d2 = [&](int a, int b) {
          double s = 0.0;
          for (int k = 0; k < 3; ++k) s += (g.X[3 * a + k] - g.X[3 * b + k]) * (g.X[3 * a + k] - g.X[3 * b + k]);
          return s;
        };
        // the octahedron around each diagonal: its other four midpoints as a ring
        const int diag[3][2] = {{m01, m23}, {m02, m13}, {m03, m12}};
        const int ring[3][4] = {{m02, m03, m13, m12}, {m01, m03, m23, m12}, {m01, m02, m23, m13}};
        int best = 0;
        for (int k = 1; k < 3; ++k)
          if (d2(diag[k][0], diag[k][1]) < d2(diag[best][0], diag[best][1])) best = k;
        for (int k = 0; k < 4; ++k) child({diag[best][0], diag[best][1], ring[best][k], ring[best][(k + 1) % 4]});
      }
    }
    g.conn.swap(conn);
    g.gmsh_id.swap(gid);
    g.nvert = static_cast<long>(g.X.size() / 3);
    g.ncell = static_cast<long>(g.gmsh_id.size());
    if (g.nvert >= (1L << 31) / 2) throw NotImplemented("grid too large for 32-bit slot indices");
  }
}

void Grid::cell_ijk(long cell, long ijk[3]) const {
  ijk[0] = cell % n[0];
  long r = cell / n[0];
  ijk[1] = dim > 1 ? r % n[1] : 0;
  ijk[2] = dim > 2 ? r / n[1] : 0;
}

long Grid::cell_vertex(long cell, int c) const {
  if (unstructured) return conn[static_cast<size_t>(cell) * nvpe + c];
  long ijk[3];
  cell_ijk(cell, ijk);
  return vertex(ijk[0] + (c & 1), ijk[1] + ((c >> 1) & 1), ijk[2] + ((c >> 2) & 1));
}

void Grid::vertex_pos(long v, double x[3]) const {
  if (unstructured) {
    for (int a = 0; a < 3; ++a) x[a] = a < dim ? X[3 * static_cast<size_t>(v) + a] : 0.0;
    return;
  }
  long i = v % (n[0] + 1);
  long r = v / (n[0] + 1);
  long j = dim > 1 ? r % (n[1] + 1) : 0;
  long k = dim > 2 ? r / (n[1] + 1) : 0;
  long idx[3] = {i, j, k};
  for (int a = 0; a < 3; ++a) x[a] = a < dim ? origin[a] + h[a] * static_cast<double>(idx[a]) : 0.0;
  // exact end points, as the structured factory places them
  for (int a = 0; a < dim; ++a)
    if (idx[a] == n[a]) x[a] = origin[a] + h[a] * static_cast<double>(n[a]);
}

// ------------------------------------------------------------------ rules
Rule element_rule(int dim, bool cube) {
  Rule r;
  const double g0 = 0.5 - 0.5 / std::sqrt(3.0), g1 = 0.5 + 0.5 / std::sqrt(3.0);
  if (dim == 0) {
    r.x.push_back({0, 0, 0});
    r.w.push_back(1.0);
    return r;
  }
  if (cube || dim == 1) {
    int n = 1 << dim;
    for (int i = 0; i < n; ++i) {
      std::array<double, 3> p{0, 0, 0};
      for (int k = 0; k < dim; ++k) p[k] = ((i >> k) & 1) ? g1 : g0;
      r.x.push_back(p);
      r.w.push_back(1.0 / n);
    }
    return r;
  }
  if (dim == 2) {
    // 3-point interior rule of order 2
    r.x = {{1.0 / 6, 1.0 / 6, 0}, {2.0 / 3, 1.0 / 6, 0}, {1.0 / 6, 2.0 / 3, 0}};
    r.w = {1.0 / 6, 1.0 / 6, 1.0 / 6};
    return r;
  }
  // 4-point tetrahedron rule of order 2
  const double a = (5.0 + 3.0 * std::sqrt(5.0)) / 20.0, b = (5.0 - std::sqrt(5.0)) / 20.0;
  r.x = {{b, b, b}, {a, b, b}, {b, a, b}, {b, b, a}};
  r.w = {1.0 / 24, 1.0 / 24, 1.0 / 24, 1.0 / 24};
  return r;
}

Rule facet_rule(int dim, bool cube) { return element_rule(dim - 1, cube); }

Rule reduce_rule(int dim, bool cube) {
  Rule r;
  if (dim == 0) {
    r.x.push_back({0, 0, 0});
    r.w.push_back(1.0);
    return r;
  }
  if (cube || dim == 1) {
    const double s = 0.5 * std::sqrt(0.6);
    const double x1[3] = {0.5 - s, 0.5, 0.5 + s}, w1[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};
    int n = 1;
    for (int k = 0; k < dim; ++k) n *= 3;
    for (int i = 0; i < n; ++i) {
      std::array<double, 3> p{0, 0, 0};
      double w = 1.0;
      for (int k = 0, m = i; k < dim; ++k, m /= 3) {
        p[k] = x1[m % 3];
        w *= w1[m % 3];
      }
      r.x.push_back(p);
      r.w.push_back(w);
    }
    return r;
  }
  if (dim == 2) {
    // two orbits of three points (area coordinates (a, a, 1 - 2a))
    const double a[2] = {0.445948490915965, 0.091576213509771};
    const double w[2] = {0.223381589678011 / 2.0, 0.109951743655322 / 2.0};
    for (int o = 0; o < 2; ++o) {
      const double b = 1.0 - 2.0 * a[o];
      r.x.push_back({a[o], a[o], 0});
      r.x.push_back({b, a[o], 0});
      r.x.push_back({a[o], b, 0});
      for (int k = 0; k < 3; ++k) r.w.push_back(w[o]);
    }
    return r;
  }
  // centroid, an orbit of 4 (a, a, a, 1 - 3a) and an orbit of 6 (b, b, c, c)
  r.x.push_back({0.25, 0.25, 0.25});
  r.w.push_back(-74.0 / 5625.0);
  const double a = 1.0 / 14.0, a3 = 1.0 - 3.0 * a;
  for (auto p : std::vector<std::array<double, 3>>{{a, a, a}, {a3, a, a}, {a, a3, a}, {a, a, a3}}) {
    r.x.push_back(p);
    r.w.push_back(343.0 / 45000.0);
  }
  const double b = 0.25 * (1.0 + std::sqrt(5.0 / 14.0)), c = 0.25 * (1.0 - std::sqrt(5.0 / 14.0));
  for (auto p : std::vector<std::array<double, 3>>{{b, b, c}, {b, c, b}, {c, b, b}, {c, c, b}, {c, b, c}, {b, c, c}}) {
    r.x.push_back(p);
    r.w.push_back(56.0 / 2250.0);
  }
  return r;
}

// ------------------------------------------------------------------ marking
std::vector<int8_t> mark_elements(const Grid& g, const std::vector<expr::Ex>& comp_expr) {
  if (comp_expr.size() > 120) throw NotImplemented("more than 120 compartments");
  std::map<std::string, int> slots = {{"position_x", 0}, {"position_y", 1}, {"position_z", 2}};
  if (g.unstructured) slots["gmsh_id"] = 3;  // cell data of gmsh grids (make_multi_domain_grid.hh:99-101)
  std::vector<expr::Bound> ev;
  std::vector<char> null;
  for (auto& e : comp_expr) {
    null.push_back(e ? 0 : 1);
    ev.push_back(e ? expr::bind(e, slots) : expr::Bound{});
  }
  std::vector<int8_t> out(static_cast<size_t>(g.ncell) * g.nclass, -1);
  for (long cell = 0; cell < g.ncell; ++cell) {
    double cx[8][3];
    for (int c = 0; c < g.ncorner; ++c) g.vertex_pos(g.cell_vertex(cell, c), cx[c]);
    for (int cls = 0; cls < g.nclass; ++cls) {
      double x[4] = {0, 0, 0, g.unstructured ? g.gmsh_id[cell] : 0.0};
      for (int a = 0; a < g.nvpe; ++a)
        for (int k = 0; k < 3; ++k) x[k] += cx[g.corner[cls][a]][k] / g.nvpe;
      int found = -1;
      for (size_t i = 0; i < ev.size(); ++i) {
        if (null[i]) continue;
        if (ev[i](x) != 0.0) {
          if (found >= 0) throw NotImplemented("overlapping compartments are not supported");
          found = static_cast<int>(i);
        }
      }
      out[static_cast<size_t>(cell) * g.nclass + cls] = static_cast<int8_t>(found);
    }
  }
  return out;
}

// ------------------------------------------------------------------ layout
static int stencil_index(int dim, const int* delta) {
  int o = 0, p = 1;
  for (int k = 0; k < dim; ++k) {
    o += (delta[k] + 1) * p;
    p *= 3;
  }
  return o;
}

static void corner_delta(int ca, int cc, int dim, int* d) {
  for (int k = 0; k < dim; ++k) d[k] = ((cc >> k) & 1) - ((ca >> k) & 1);
}

namespace {
struct FaceKey {
  long v[4];
  bool operator==(const FaceKey& o) const { return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2] && v[3] == o.v[3]; }
};
struct FaceHash {
  size_t operator()(const FaceKey& k) const {
    size_t h = 1469598103934665603ull;
    for (long x : k.v) h = (h ^ static_cast<size_t>(x)) * 1099511628211ull;
    return h;
  }
};
// local vertex lists of element faces (reference element numbering)
std::vector<std::vector<int>> element_faces(const Grid& g) {
  std::vector<std::vector<int>> f;
  if (g.cube) {
    for (int a = 0; a < g.dim; ++a)
      for (int s = 0; s < 2; ++s) {
        std::vector<int> l;
        for (int v = 0; v < g.nvpe; ++v)
          if (((v >> a) & 1) == s) l.push_back(v);
        f.push_back(l);
      }
  } else {
    for (int k = 0; k <= g.dim; ++k) {
      std::vector<int> l;
      for (int v = 0; v <= g.dim; ++v)
        if (v != k) l.push_back(v);
      f.push_back(l);
    }
  }
  return f;
}
}  // namespace

static void facet_geometry(const Grid& g, long cell, int cls, const std::vector<int>& fl, Facet& F) {
  const int d = g.dim;
  std::vector<std::array<double, 3>> P(fl.size());
  for (size_t i = 0; i < fl.size(); ++i) g.vertex_pos(g.cell_vertex(cell, g.corner[cls][fl[i]]), P[i].data());
  double t1[3] = {0, 0, 0}, t2[3] = {0, 0, 0};
  if (d >= 2)
    for (int k = 0; k < 3; ++k) t1[k] = P[1][k] - P[0][k];
  if (d == 3)
    for (int k = 0; k < 3; ++k) t2[k] = P[2][k] - P[0][k];
  double ie = 1.0, n[3] = {0, 0, 0};
  if (d == 1) {
    n[0] = 1.0;
  } else if (d == 2) {
    ie = std::hypot(t1[0], t1[1]);
    n[0] = t1[1] / ie;
    n[1] = -t1[0] / ie;
  } else {
    double c[3] = {t1[1] * t2[2] - t1[2] * t2[1], t1[2] * t2[0] - t1[0] * t2[2], t1[0] * t2[1] - t1[1] * t2[0]};
    ie = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    for (int k = 0; k < 3; ++k) n[k] = c[k] / ie;
  }
  // orient away from the inside element
  double ec[3] = {0, 0, 0}, fc[3] = {0, 0, 0};
  for (int a = 0; a < g.nvpe; ++a) {
    double x[3];
    g.vertex_pos(g.cell_vertex(cell, g.corner[cls][a]), x);
    for (int k = 0; k < 3; ++k) ec[k] += x[k] / g.nvpe;
  }
  for (auto& p : P)
    for (int k = 0; k < 3; ++k) fc[k] += p[k] / static_cast<double>(P.size());
  double s = 0;
  for (int k = 0; k < 3; ++k) s += n[k] * (fc[k] - ec[k]);
  if (s < 0)
    for (double& v : n) v = -v;
  Rule r = facet_rule(d, g.cube);
  F.volume = 0;
  const int nvf = static_cast<int>(fl.size());
  for (size_t q = 0; q < r.w.size(); ++q) {
    double eta0 = r.x[q][0], eta1 = r.x[q][1];
    for (int k = 0; k < 3; ++k) F.qx.push_back(P[0][k] + eta0 * t1[k] + eta1 * t2[k]);
    for (int k = 0; k < 3; ++k) F.qn.push_back(n[k]);
    F.qw.push_back(r.w[q] * ie);
    F.volume += r.w[q] * ie;
    for (int f = 0; f < nvf; ++f) {
      double phi;
      if (d == 1) phi = 1.0;
      else if (g.cube) {  // tensor basis on the face, vertices in lexicographic order
        phi = ((f & 1) ? eta0 : 1.0 - eta0);
        if (d == 3) phi *= ((f >> 1) & 1) ? eta1 : 1.0 - eta1;
      } else {  // barycentric on the simplex face
        phi = f == 0 ? 1.0 - eta0 - (d == 3 ? eta1 : 0.0) : (f == 1 ? eta0 : eta1);
      }
      F.phi.push_back(phi);
    }
  }
}

Layout make_layout(const Grid& g, const std::vector<int8_t>& elem_comp, std::vector<Compartment> comps,
                   bool boundary_facets, bool interface_facets) {
  Layout L;
  L.grid = g;
  const int d = g.dim;
  const int nstencil = d == 1 ? 3 : d == 2 ? 9 : 27;
  L.B = 1;
  for (auto& C : comps) L.B = std::max(L.B, C.b);
  if (L.B > 8) throw NotImplemented("more than 8 species in one compartment");
  long slot = 0, dof = 0;
  for (size_t ci = 0; ci < comps.size(); ++ci) {
    Compartment& C = comps[ci];
    C.g2l.assign(g.nvert, -1);
    C.nodes.clear();
    C.cells.clear();
    C.cell_mask.clear();
    if (C.b > 0) {
      std::vector<char> used(g.nvert, 0);
      std::vector<uint8_t> mask(g.ncell, 0);
      for (long cell = 0; cell < g.ncell; ++cell)
        for (int cls = 0; cls < g.nclass; ++cls)
          if (elem_comp[static_cast<size_t>(cell) * g.nclass + cls] == static_cast<int8_t>(ci)) {
            mask[cell] |= static_cast<uint8_t>(1u << cls);
            for (int a = 0; a < g.nvpe; ++a) used[g.cell_vertex(cell, g.corner[cls][a])] = 1;
          }
      for (long v = 0; v < g.nvert; ++v)
        if (used[v]) {
          C.g2l[v] = static_cast<int>(C.nodes.size());
          C.nodes.push_back(v);
        }
      // colour = parity pattern of the cell index: same-colour cells share no
      // vertex; unstructured cells are coloured greedily in cell order
      std::vector<std::vector<int>> by_color(g.unstructured ? 0 : 1 << d);
      std::vector<uint64_t> vcol(g.unstructured ? g.nvert : 0, 0);
      for (long cell = 0; cell < g.ncell; ++cell) {
        if (!mask[cell]) continue;
        int col = 0;
        if (g.unstructured) {
          uint64_t taken = 0;
          for (int a = 0; a < g.nvpe; ++a) taken |= vcol[g.cell_vertex(cell, a)];
          while (col < 64 && ((taken >> col) & 1)) ++col;
          if (col == 64) throw Error("cell colouring overflow");
          for (int a = 0; a < g.nvpe; ++a) vcol[g.cell_vertex(cell, a)] |= 1ull << col;
          if (static_cast<int>(by_color.size()) <= col) by_color.resize(col + 1);
        } else {
          long ijk[3];
          g.cell_ijk(cell, ijk);
          for (int k = 0; k < d; ++k) col |= static_cast<int>(ijk[k] & 1) << k;
        }
        by_color[col].push_back(static_cast<int>(cell));
      }
      C.color_ptr.assign(1, 0);
      for (auto& lst : by_color) {
        for (int cell : lst) {
          C.cells.push_back(cell);
          C.cell_mask.push_back(mask[cell]);
        }
        C.color_ptr.push_back(static_cast<int>(C.cells.size()));
      }
    } else {
      C.color_ptr.assign(1, 0);
    }
    C.nslot = static_cast<long>(C.nodes.size());
    C.slot_off = slot;
    C.dof_off = dof;
    slot += C.nslot;
    dof += C.nslot * C.b;
  }
  L.nslot = slot;
  L.ndof = dof;
  L.slot_comp.assign(slot, -1);
  for (size_t ci = 0; ci < comps.size(); ++ci)
    for (long s = 0; s < comps[ci].nslot; ++s) L.slot_comp[comps[ci].slot_off + s] = static_cast<int>(ci);
  // cell corner slots and own-compartment stencil masks (unstructured: the
  // own-compartment neighbour lists)
  L.own_mask.assign(slot, 0u);
  std::vector<std::vector<int>> own_nb(g.unstructured ? slot : 0);
  for (auto& C : comps) {
    C.cell_slots.assign(C.cells.size() * g.ncorner, -1);
    for (size_t i = 0; i < C.cells.size(); ++i) {
      long cell = C.cells[i];
      for (int c = 0; c < g.ncorner; ++c) {
        int l = C.g2l[g.cell_vertex(cell, c)];
        C.cell_slots[i * g.ncorner + c] = l < 0 ? -1 : static_cast<int>(C.slot_off + l);
      }
      for (int cls = 0; cls < g.nclass; ++cls) {
        if (!((C.cell_mask[i] >> cls) & 1)) continue;
        for (int a = 0; a < g.nvpe; ++a)
          for (int b = 0; b < g.nvpe; ++b) {
            int ca = g.corner[cls][a], cb = g.corner[cls][b];
            int sa = C.cell_slots[i * g.ncorner + ca];
            if (g.unstructured) {
              own_nb[sa].push_back(C.cell_slots[i * g.ncorner + cb]);
              continue;
            }
            int delta[3];
            corner_delta(ca, cb, d, delta);
            L.own_mask[sa] |= 1u << stencil_index(d, delta);
          }
      }
    }
  }
  // facets (outflow terms only)
  std::vector<std::vector<int>> extra(slot);
  if (boundary_facets || interface_facets) {
    auto faces = element_faces(g);
    std::unordered_map<FaceKey, std::pair<long, int>, FaceHash> open;
    open.reserve(1024);
    struct Pair {
      long e1;
      int f1;
      long e2;
    };
    std::vector<Pair> pairs;
    // only elements that belong to some compartment and touch a compartment
    // change or the boundary matter; hashing every face keeps this generic.
    for (long cell = 0; cell < g.ncell; ++cell)
      for (int cls = 0; cls < g.nclass; ++cls) {
        long e = cell * g.nclass + cls;
        for (size_t f = 0; f < faces.size(); ++f) {
          FaceKey key{{-1, -1, -1, -1}};
          int nvf = static_cast<int>(faces[f].size());
          for (int i = 0; i < nvf; ++i) key.v[i] = g.cell_vertex(cell, g.corner[cls][faces[f][i]]);
          std::sort(key.v, key.v + nvf);
          auto it = open.find(key);
          if (it == open.end()) {
            open.emplace(key, std::make_pair(e, static_cast<int>(f)));
          } else {
            int c1 = elem_comp[it->second.first], c2 = elem_comp[e];
            if (c1 != c2 && c1 >= 0 && c2 >= 0) pairs.push_back({it->second.first, it->second.second, e});
            open.erase(it);
          }
        }
      }
    if (boundary_facets)
      for (auto& kv : open)
        if (elem_comp[kv.second.first] >= 0) pairs.push_back({kv.second.first, kv.second.second, -1});
    if (!interface_facets) {
      std::vector<Pair> keep;
      for (auto& p : pairs)
        if (p.e2 < 0) keep.push_back(p);
      pairs.swap(keep);
    }
    std::sort(pairs.begin(), pairs.end(), [](const Pair& a, const Pair& b) {
      return a.e1 != b.e1 ? a.e1 < b.e1 : a.f1 < b.f1;
    });
    std::vector<Facet> raw;
    for (auto& p : pairs) {
      Facet F;
      long cell = p.e1 / g.nclass;
      int cls = static_cast<int>(p.e1 % g.nclass);
      F.comp_in = elem_comp[p.e1];
      F.boundary = p.e2 < 0;
      F.comp_out = F.boundary ? F.comp_in : elem_comp[p.e2];
      const auto& fl = faces[p.f1];
      for (int lv : fl) {
        long v = g.cell_vertex(cell, g.corner[cls][lv]);
        const Compartment& Ci = comps[F.comp_in];
        const Compartment& Co = comps[F.comp_out];
        F.slot_in.push_back(Ci.b ? static_cast<int>(Ci.slot_off + Ci.g2l[v]) : -1);
        F.slot_out.push_back(Co.b ? static_cast<int>(Co.slot_off + Co.g2l[v]) : -1);
      }
      facet_geometry(g, cell, cls, fl, F);
      if (!F.boundary && comps[F.comp_in].b && comps[F.comp_out].b)
        for (int si : F.slot_in)
          for (int so : F.slot_out) {
            extra[si].push_back(so);
            extra[so].push_back(si);
          }
      raw.push_back(std::move(F));
    }
    // greedy colouring: facets of one colour touch disjoint slots
    std::vector<uint64_t> used(slot, 0);
    std::vector<std::vector<Facet>> colored;
    for (auto& F : raw) {
      uint64_t m = 0;
      for (int s : F.slot_in)
        if (s >= 0) m |= used[s];
      for (int s : F.slot_out)
        if (s >= 0) m |= used[s];
      int c = 0;
      while (c < 64 && ((m >> c) & 1)) ++c;
      if (c == 64) throw Error("facet colouring overflow");
      if (static_cast<int>(colored.size()) <= c) colored.resize(c + 1);
      for (int s : F.slot_in)
        if (s >= 0) used[s] |= 1ull << c;
      for (int s : F.slot_out)
        if (s >= 0) used[s] |= 1ull << c;
      colored[c].push_back(std::move(F));
    }
    L.facet_color_ptr.assign(1, 0);
    for (auto& lst : colored) {
      for (auto& F : lst) L.facets.push_back(std::move(F));
      L.facet_color_ptr.push_back(static_cast<int>(L.facets.size()));
    }
  } else {
    L.facet_color_ptr.assign(1, 0);
  }
  // rows: [lower-compartment extras][own stencil][higher-compartment extras]
  const long sx = 1, sy = g.n[0] + 1, sz = (g.n[0] + 1) * (g.n[1] + 1);
  L.row_ptr.assign(slot + 1, 0);
  L.own_start.assign(slot, 0);
  L.diag_pos.assign(slot, -1);
  long nnz = 0;
  for (long s = 0; s < slot; ++s) {
    auto& ex = extra[s];
    std::sort(ex.begin(), ex.end());
    ex.erase(std::unique(ex.begin(), ex.end()), ex.end());
    if (g.unstructured) {
      auto& nb = own_nb[s];
      std::sort(nb.begin(), nb.end());
      nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
      nnz += static_cast<long>(ex.size() + nb.size());
    } else {
      nnz += static_cast<long>(ex.size()) + __builtin_popcount(L.own_mask[s]);
    }
  }
  if (nnz >= (1L << 31)) throw NotImplemented("BCSR pattern exceeds 32-bit indexing");
  L.col.resize(nnz);
  long p = 0;
  for (size_t ci = 0; ci < comps.size(); ++ci) {
    const Compartment& C = comps[ci];
    for (long l = 0; l < C.nslot; ++l) {
      long s = C.slot_off + l;
      L.row_ptr[s] = static_cast<int>(p);
      auto& ex = extra[s];
      size_t e = 0;
      while (e < ex.size() && ex[e] < C.slot_off) L.col[p++] = ex[e++];
      L.own_start[s] = static_cast<int>(p);
      long v = C.nodes[l];
      if (g.unstructured) {
        for (int t : own_nb[s]) {
          if (t == s) L.diag_pos[s] = static_cast<int>(p);
          L.col[p++] = t;
        }
        while (e < ex.size()) L.col[p++] = ex[e++];
        continue;
      }
      for (int o = 0; o < nstencil; ++o) {
        if (!((L.own_mask[s] >> o) & 1)) continue;
        int dx = o % 3 - 1, dy = d > 1 ? (o / 3) % 3 - 1 : 0, dz = d > 2 ? o / 9 - 1 : 0;
        long w = v + dx * sx + dy * sy + dz * sz;
        int t = C.g2l[w];
        if (t < 0) throw Error("internal: stencil neighbour outside compartment");
        if (o == nstencil / 2) L.diag_pos[s] = static_cast<int>(p);
        L.col[p++] = static_cast<int>(C.slot_off + t);
      }
      while (e < ex.size()) L.col[p++] = ex[e++];
    }
  }
  L.row_ptr[slot] = static_cast<int>(p);
  for (long s = 0; s < slot; ++s)
    if (L.diag_pos[s] < 0) throw Error("internal: missing diagonal block");
  if (g.unstructured) {
    // block of every (a, c) vertex pair of every cell, for the cell kernels
    for (auto& C : comps) {
      const int nv = g.nvpe;
      C.cell_pos.assign(C.cells.size() * nv * nv, -1);
      for (size_t i = 0; i < C.cells.size(); ++i)
        for (int a = 0; a < nv; ++a)
          for (int c = 0; c < nv; ++c) {
            const int sa = C.cell_slots[i * g.ncorner + a], sc = C.cell_slots[i * g.ncorner + c];
            const int* b = L.col.data() + L.own_start[sa];
            const int* e = b + own_nb[sa].size();
            const int* it = std::lower_bound(b, e, sc);
            if (it == e || *it != sc) throw Error("internal: cell block missing from pattern");
            C.cell_pos[(i * nv + a) * nv + c] = static_cast<int>(it - L.col.data());
          }
    }
  }
  L.comps = std::move(comps);
  return L;
}

}  // namespace dcopasi
