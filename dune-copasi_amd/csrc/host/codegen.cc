#include "codegen.hh"

#include "../kernels/kernels.hh"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <sstream>

#include "error.hh"
#include "reduce.hh"

namespace dcopasi {

namespace {

std::string D(double v) {
  char b[40];
  std::snprintf(b, sizeof b, "%.17g", v);
  std::string s = b;
  if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
  return s;
}

const char* kAxis = "xyz";

// reference shape functions
void shape(int dim, bool cube, const double* xi, double* phi, double (*dphi)[3]) {
  if (cube) {
    int nv = 1 << dim;
    for (int v = 0; v < nv; ++v) {
      double p = 1.0, g[3] = {1, 1, 1};
      for (int a = 0; a < dim; ++a) {
        bool b = (v >> a) & 1;
        double f = b ? xi[a] : 1.0 - xi[a];
        for (int c = 0; c < dim; ++c) g[c] *= (c == a) ? (b ? 1.0 : -1.0) : f;
        p *= f;
      }
      phi[v] = p;
      for (int c = 0; c < 3; ++c) dphi[v][c] = c < dim ? g[c] : 0.0;
    }
  } else {
    double s = 0;
    for (int a = 0; a < dim; ++a) s += xi[a];
    phi[0] = 1.0 - s;
    for (int c = 0; c < 3; ++c) dphi[0][c] = c < dim ? -1.0 : 0.0;
    for (int v = 1; v <= dim; ++v) {
      phi[v] = xi[v - 1];
      for (int c = 0; c < 3; ++c) dphi[v][c] = (c == v - 1) ? 1.0 : 0.0;
    }
  }
}

void invert(int d, double A[3][3], double B[3][3], double& det) {
  if (d == 1) {
    det = A[0][0];
    B[0][0] = 1.0 / det;
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

struct ElementTables {
  int nq, nv;
  std::vector<double> phi;   // [q][a]
  std::vector<double> gphi;  // [p][q][a][k] physical gradients
  std::vector<double> fac;   // [q]
  std::vector<double> xq;    // [p][q][3] offset of the point from the cell origin
  std::vector<int> off;      // [p][a][c] stencil index
  double evol;
};

ElementTables element_tables(const Grid& g, const Rule* rule = nullptr) {
  ElementTables T;
  Rule r = rule ? *rule : element_rule(g.dim, g.cube);
  T.nq = static_cast<int>(r.w.size());
  T.nv = g.nvpe;
  const int d = g.dim;
  double phi[8], dref[8][3];
  for (int q = 0; q < T.nq; ++q) {
    shape(d, g.cube, r.x[q].data(), phi, dref);
    for (int a = 0; a < T.nv; ++a) T.phi.push_back(phi[a]);
  }
  double det = 0;
  for (int p = 0; p < g.nclass; ++p) {
    // affine map of class p: x = cell_origin + J xi, J columns from the corners
    double J[3][3] = {{0}}, Ji[3][3] = {{0}};
    auto cpos = [&](int corner, int k) { return ((corner >> k) & 1) ? g.h[k] : 0.0; };
    for (int col = 0; col < d; ++col) {
      int cv = g.cube ? (1 << col) : g.corner[p][col + 1];
      int c0 = g.corner[p][0];
      // unstructured cells: the reference simplex (J = I); the element's own
      // geometry comes at run time
      for (int k = 0; k < d; ++k) J[k][col] = g.unstructured ? (k == col ? 1.0 : 0.0) : cpos(cv, k) - cpos(c0, k);
    }
    invert(d, J, Ji, det);
    for (int q = 0; q < T.nq; ++q) {
      shape(d, g.cube, r.x[q].data(), phi, dref);
      for (int a = 0; a < T.nv; ++a)
        for (int k = 0; k < d; ++k) {
          double s = 0;
          for (int l = 0; l < d; ++l) s += dref[a][l] * Ji[l][k];
          T.gphi.push_back(s);
        }
      for (int k = 0; k < 3; ++k) {
        double s = k < d && !g.unstructured ? cpos(g.corner[p][0], k) : 0.0;
        for (int l = 0; l < d && k < d; ++l) s += J[k][l] * r.x[q][l];
        T.xq.push_back(s);
      }
    }
    for (int a = 0; a < T.nv; ++a)
      for (int c = 0; c < T.nv; ++c) {
        int ca = g.corner[p][a], cc = g.corner[p][c];
        int o = 0, m = 1;
        for (int k = 0; k < d; ++k) {
          o += (((cc >> k) & 1) - ((ca >> k) & 1) + 1) * m;
          m *= 3;
        }
        T.off.push_back(o);
      }
  }
  T.evol = 0;
  for (int q = 0; q < T.nq; ++q) {
    T.fac.push_back(r.w[q] * det);
    T.evol += r.w[q] * det;
  }
  return T;
}

template <class V>
std::string array_init(const V& v) {
  std::ostringstream s;
  s << "{";
  for (size_t i = 0; i < v.size(); ++i) s << (i ? "," : "") << D(static_cast<double>(v[i]));
  s << "}";
  return s.str();
}
std::string int_init(const std::vector<int>& v) {
  std::ostringstream s;
  s << "{";
  for (size_t i = 0; i < v.size(); ++i) s << (i ? "," : "") << v[i];
  s << "}";
  return s.str();
}

bool tensor_const(const TensorTerm& t) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (t.d[i][j] && !expr::is_const(t.d[i][j])) return false;
  return true;
}

// diffusion whose entries depend on nothing but run-time parameters and time:
// its stiffness contribution is D times a reference table
bool tensor_uniform(const TensorTerm& t) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      if (!t.d[i][j]) continue;
      std::vector<std::string> vars;
      expr::collect_vars(t.d[i][j], vars);
      for (const auto& v : vars)
        if (v != "time" && v.rfind("$", 0) != 0) return false;
    }
  return true;
}

// (i, j) entries of the point coefficient block a compartment's point
// function writes: storage diagonal, storage and reaction Jacobians
std::vector<std::pair<int, int>> mc_pattern(const ModelSpec& m, int comp) {
  std::vector<std::pair<int, int>> nz;
  auto add = [&](int i, int j) {
    if (std::find(nz.begin(), nz.end(), std::make_pair(i, j)) == nz.end()) nz.push_back({i, j});
  };
  const auto& species = m.comp_species[comp];
  for (size_t i = 0; i < species.size(); ++i) {
    const auto& sp = m.species[species[i]];
    const int I = static_cast<int>(i);
    if (sp.storage) {
      add(I, I);
      for (const auto& pj : sp.storage_jac)
        if (m.species[pj.wrt].comp == comp) add(I, m.species[pj.wrt].local);
    }
    if (sp.reaction)
      for (const auto& pj : sp.reaction_jac)
        if (m.species[pj.wrt].comp == comp) add(I, m.species[pj.wrt].local);
  }
  std::sort(nz.begin(), nz.end());
  return nz;
}

const char* kPrelude = R"(
#define DCG_B %B%
// DCG_GEO: element tables are those of the unit cell and the spacing comes at
// run time in prm[0..7] = {h_x, h_y, h_z, 1/h_x, 1/h_y, 1/h_z, |cell|, 0}
#define DCG_GEO %GEO%
// DCG_UNSTRUCTURED: every element carries its own geometry in eg[]: the
// gradients of its basis functions, the points and |det J| (UG_* offsets)
#define DCG_UNSTRUCTURED %UNS%
#if DCG_UNSTRUCTURED
#define GP(p, q, a, k) (eg[(a) * DIM + (k)])
#define FACQ(q) (FAC[q] * eg[UG_DET])
#define XQP(p, q, k) (eg[UG_POS + (q) * 3 + (k)])
#elif DCG_GEO
#define GP(p, q, a, k) (GPHI[p][q][a][k] * prm[3 + (k)])
#define FACQ(q) (FAC[q] * prm[6])
#define XQP(p, q, k) (XQ[p][q][k] * prm[k])
#else
#define GP(p, q, a, k) GPHI[p][q][a][k]
#define FACQ(q) FAC[q]
#define XQP(p, q, k) XQ[p][q][k]
#endif
// descriptor of one batch member for the batched tile kernels
struct dcg_asm_member {
  double* vec[8];
  double* val;
  const int* own_start;
  const unsigned* own_mask;
  const double* prm;
  const int* g2l[4];
  const unsigned char* cmask[4];
  int n[4];
  double origin[3];
  double h[3];
  double* pc[4];
  double* pdiag;
  double* tca;
};
// descriptor of one batch member for the batched point-coefficient apply
struct dcg_pc_member {
  const int* g2l;
  const unsigned char* cmask;
  const double* pc;
  const float* pc32;
  const double* tca;
  const double* prm;
  int n[4];
  double* v[17];
  float* f[10];
  const int* done;
  const int* irdone;
};
static_assert(sizeof(dcg_asm_member) == %ASM_BYTES%, "batch member descriptor layout");
static_assert(sizeof(dcg_pc_member) == %PCM_BYTES%, "point member descriptor layout");
__device__ __forceinline__ double dcg_sign(double v) { return (double)((v > 0.0) - (v < 0.0)); }
__device__ __forceinline__ int dcg_pos(const int* own_start, const unsigned* own_mask, int s, int o) {
  return own_start[s] + __popc(own_mask[s] & ((1u << o) - 1u));
}
)";

// Volume kernel body; %C% is the compartment namespace.
const char* kVolume = R"(
namespace %C% {
template <int P>
__device__ __forceinline__ void residual(const double (&xl)[NV][BC], double cx, double cy, double cz, double t,
                                         double cm, double ca, double (&rl)[NV][BC], double (*mc)[BC][BC],
                                         double (*gc)[BC][BC][DIM], double (*dc)[BC][BC][DIM][DIM],
                                         const double* __restrict__ prm, const double* __restrict__ eg) {
#pragma unroll
  for (int a = 0; a < NV; ++a)
#pragma unroll
    for (int i = 0; i < BC; ++i) rl[a][i] = 0.0;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double u[BC], g[BC][3];
#pragma unroll
    for (int i = 0; i < BC; ++i) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < NV; ++a) s += PHI[q][a] * xl[a][i];
      u[i] = s;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double gs = 0.0;
        if (k < DIM) {
#pragma unroll
          for (int a = 0; a < NV; ++a) gs += GP(P, q, a, k < DIM ? k : 0) * xl[a][i];
        }
        g[i][k] = gs;
      }
    }
    const double px = cx + XQP(P, q, 0), py = cy + XQP(P, q, 1), pz = cz + XQP(P, q, 2);
    (void)px; (void)py; (void)pz;
    double rs[BC], rf[BC][DIM];
    qp(q, u, g, px, py, pz, t, cm, ca, rs, rf, mc ? mc[q] : nullptr, gc ? gc[q] : nullptr, dc ? dc[q] : nullptr, prm,
       eg);
#pragma unroll
    for (int a = 0; a < NV; ++a)
#pragma unroll
      for (int i = 0; i < BC; ++i) {
        double s = rs[i] * PHI[q][a];
#pragma unroll
        for (int k = 0; k < DIM; ++k) s += rf[i][k] * GP(P, q, a, k);
        rl[a][i] += s;
      }
  }
}

template <int P>
__device__ __forceinline__ void element(const double (&xc)[NCORNER][BC], const int (&sl)[NCORNER], double cx,
                                        double cy, double cz, double t, double cm, double ca, double* r,
                                        double* val, const int* own_start, const unsigned* own_mask, int mode,
                                        const double* __restrict__ prm, const double* __restrict__ eg) {
  double xl[NV][BC];
#pragma unroll
  for (int a = 0; a < NV; ++a)
#pragma unroll
    for (int i = 0; i < BC; ++i) xl[a][i] = xc[CORNER[P][a]][i];
  double rl[NV][BC];
  const bool want_j = (mode & 2) != 0;
#if NUMERICAL_JAC
  residual<P>(xl, cx, cy, cz, t, cm, ca, rl, nullptr, nullptr, nullptr, prm, eg);
  if (want_j) {
#pragma unroll 1
    for (int c = 0; c < NV; ++c) {
#pragma unroll 1
      for (int j = 0; j < BC; ++j) {
        double xp[NV][BC];
#pragma unroll
        for (int a = 0; a < NV; ++a)
#pragma unroll
          for (int i = 0; i < BC; ++i) xp[a][i] = xl[a][i];
        const double delta = FD_EPS * (1.0 + fabs(xl[c][j]));
        xp[c][j] += delta;
        double rp[NV][BC];
        residual<P>(xp, cx, cy, cz, t, cm, ca, rp, nullptr, nullptr, nullptr, prm, eg);
#pragma unroll
        for (int a = 0; a < NV; ++a) {
          const int sa = sl[CORNER[P][a]];
#if DCG_UNSTRUCTURED
          const int pos = own_start[a * NV + c];
#else
          const int pos = dcg_pos(own_start, own_mask, sa, OFF[P][a][c]);
#endif
#pragma unroll
          for (int i = 0; i < BC; ++i)
            val[(size_t)pos * (DCG_B * DCG_B) + i * DCG_B + j] += (rp[a][i] - rl[a][i]) / delta;
        }
      }
    }
  }
#else
  double mc[NQ][BC][BC];
#if HAS_GC
  double gc[NQ][BC][BC][DIM];
#endif
#if HAS_DC
  double dc[NQ][BC][BC][DIM][DIM];
#endif
  residual<P>(xl, cx, cy, cz, t, cm, ca, rl, want_j ? mc : nullptr,
#if HAS_GC
              want_j ? gc : nullptr,
#else
              nullptr,
#endif
#if HAS_DC
              want_j ? dc : nullptr,
#else
              nullptr,
#endif
              prm, eg);
  if (want_j) {
#pragma unroll
    for (int a = 0; a < NV; ++a) {
      const int sa = sl[CORNER[P][a]];
      const int st = own_start[sa];
      const unsigned mk = own_mask[sa];
#pragma unroll
      for (int c = 0; c < NV; ++c) {
#if DCG_UNSTRUCTURED
        const int pos = own_start[a * NV + c];
        (void)st;
        (void)mk;
#else
        const int pos = st + __popc(mk & ((1u << OFF[P][a][c]) - 1u));
#endif
        double* blk = val + (size_t)pos * (DCG_B * DCG_B);
#pragma unroll
        for (int i = 0; i < BC; ++i)
#pragma unroll
          for (int j = 0; j < BC; ++j) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < NQ; ++q) s += PHI[q][a] * PHI[q][c] * mc[q][i][j];
#if HAS_GC
#pragma unroll
            for (int q = 0; q < NQ; ++q)
#pragma unroll
              for (int k = 0; k < DIM; ++k) s += PHI[q][c] * GP(P, q, a, k) * gc[q][i][j][k];
#endif
#if HAS_DC
#pragma unroll
            for (int q = 0; q < NQ; ++q)
#pragma unroll
              for (int k = 0; k < DIM; ++k)
#pragma unroll
                for (int l = 0; l < DIM; ++l) s += GP(P, q, a, k) * dc[q][i][j][k][l] * GP(P, q, c, l);
#endif
            s += ca * const_diffusion<P>(a, c, i, j);
            blk[i * DCG_B + j] += s;
          }
      }
    }
  }
#endif
  if (mode & 1) {
#pragma unroll
    for (int a = 0; a < NV; ++a) {
      const int sa = sl[CORNER[P][a]];
#pragma unroll
      for (int i = 0; i < BC; ++i) r[(size_t)sa * DCG_B + i] += rl[a][i];
    }
  }
}
}  // namespace %C%

extern "C" __global__ void __launch_bounds__(128)
%KERNEL%(const int* __restrict__ cells, const unsigned char* __restrict__ cmask, const int* __restrict__ cslots,
         int first, int count, const double* __restrict__ x, double t, double cm, double ca,
         double* __restrict__ r, double* __restrict__ val, const int* __restrict__ own_start,
         const unsigned* __restrict__ own_mask, int n0, int n1, double ox, double oy, double oz, double hx,
         double hy, double hz, int mode, const double* __restrict__ prm, const double* __restrict__ egeo) {
  using namespace %C%;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= count) return;
  const int idx = first + tid;
  const int cell = cells[idx];
  const unsigned msk = cmask[idx];
  int sl[NCORNER];
  double xc[NCORNER][BC];
#pragma unroll
  for (int c = 0; c < NCORNER; ++c) {
    sl[c] = cslots[(size_t)idx * NCORNER + c];
#pragma unroll
    for (int i = 0; i < BC; ++i) xc[c][i] = sl[c] >= 0 ? x[(size_t)sl[c] * DCG_B + i] : 0.0;
  }
#if DCG_UNSTRUCTURED
  // own_start: the cells' block positions [cell][a][c]; egeo: their geometry
  (void)cell;
  own_start += (size_t)idx * NV * NV;
  const double* eg = egeo + (size_t)idx * UG_WORDS;
  const double cx = 0.0, cy = 0.0, cz = 0.0;
#else
  const double* eg = nullptr;
  const int ci = cell % n0, cj = (cell / n0) % n1, ck = cell / (n0 * n1);
  const double cx = ox + ci * hx, cy = oy + cj * hy, cz = oz + ck * hz;
#endif
%DISPATCH%
}
)";

// Facet kernel body
const char* kFacet = R"(
namespace %F% {
__device__ __forceinline__ void facet_residual(const double (&xi)[NVF][BI > 0 ? BI : 1],
                                               const double (&xo)[NVF][BO > 0 ? BO : 1], const double* fd,
                                               double t, double ca, double (&ri)[NVF][BI > 0 ? BI : 1],
                                               double (&ro)[NVF][BO > 0 ? BO : 1], double* Jii, double* Jio,
                                               double* Joi, double* Joo, const double* __restrict__ prm) {
  (void)prm;
#pragma unroll
  for (int f = 0; f < NVF; ++f) {
#pragma unroll
    for (int i = 0; i < BI; ++i) ri[f][i] = 0.0;
#pragma unroll
    for (int i = 0; i < BO; ++i) ro[f][i] = 0.0;
  }
  const double nx = fd[NQF * 4 + 0], ny = fd[NQF * 4 + 1], nz = fd[NQF * 4 + 2], fvol = fd[NQF * 4 + 3];
  (void)nx; (void)ny; (void)nz; (void)fvol;
#pragma unroll
  for (int q = 0; q < NQF; ++q) {
    const double w = fd[q] * ca;
    const double px = fd[NQF + 3 * q], py = fd[NQF + 3 * q + 1], pz = fd[NQF + 3 * q + 2];
    (void)px; (void)py; (void)pz;
    double ui[BI > 0 ? BI : 1], uo[BO > 0 ? BO : 1];
#pragma unroll
    for (int i = 0; i < BI; ++i) {
      double s = 0.0;
#pragma unroll
      for (int f = 0; f < NVF; ++f) s += PHIF[q][f] * xi[f][i];
      ui[i] = s;
    }
#pragma unroll
    for (int i = 0; i < BO; ++i) {
      double s = 0.0;
#pragma unroll
      for (int f = 0; f < NVF; ++f) s += PHIF[q][f] * xo[f][i];
      uo[i] = s;
    }
    (void)ui; (void)uo;
%QPBODY%
  }
}

// mode 1: r += facet residual; 2: facet Jacobian blocks into val (BCSR
// positions); 4: r += J(x) dir over the facet blocks (the Jacobian apply);
// 8: the diagonal blocks of J(x) (Jii and Joo with f = g) added to
// val[slot * B * B], the rows' diagonal blocks of the point-coefficient
// Jacobian.  V is the type of r and dir (FP32 for the apply inside the
// refinement loop).
template <typename V>
__device__ __forceinline__ void facet_body(const int* __restrict__ fint, const double* __restrict__ fdbl, int first,
                                           int count, const double* __restrict__ x, double t, double ca,
                                           V* __restrict__ r, double* __restrict__ val, int mode,
                                           const double* __restrict__ prm, const V* __restrict__ dir) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= count) return;
  const int id = first + tid;
  const int* fi = fint + (size_t)id * FINT;
  const double* fd = fdbl + (size_t)id * FDBL;
  double xi[NVF][BI > 0 ? BI : 1], xo[NVF][BO > 0 ? BO : 1];
#pragma unroll
  for (int f = 0; f < NVF; ++f) {
#pragma unroll
    for (int i = 0; i < BI; ++i) xi[f][i] = x[(size_t)fi[f] * DCG_B + i];
#pragma unroll
    for (int i = 0; i < BO; ++i) xo[f][i] = x[(size_t)fi[NVF + f] * DCG_B + i];
  }
  double ri[NVF][BI > 0 ? BI : 1], ro[NVF][BO > 0 ? BO : 1];
  const bool want_j = (mode & 2) != 0;
#if NUMERICAL_JAC
  facet_residual(xi, xo, fd, t, ca, ri, ro, nullptr, nullptr, nullptr, nullptr, prm);
  if (want_j) {
    const int* pos = fi + 2 * NVF;
#pragma unroll 1
    for (int side = 0; side < 2; ++side) {
      const int bs = side == 0 ? BI : BO;
#pragma unroll 1
      for (int g = 0; g < NVF; ++g)
#pragma unroll 1
        for (int j = 0; j < bs; ++j) {
          double xi2[NVF][BI > 0 ? BI : 1], xo2[NVF][BO > 0 ? BO : 1];
#pragma unroll
          for (int f = 0; f < NVF; ++f) {
#pragma unroll
            for (int i = 0; i < BI; ++i) xi2[f][i] = xi[f][i];
#pragma unroll
            for (int i = 0; i < BO; ++i) xo2[f][i] = xo[f][i];
          }
          double& xv = side == 0 ? xi2[g][j] : xo2[g][j];
          const double delta = FD_EPS * (1.0 + fabs(xv));
          xv += delta;
          double ri2[NVF][BI > 0 ? BI : 1], ro2[NVF][BO > 0 ? BO : 1];
          facet_residual(xi2, xo2, fd, t, ca, ri2, ro2, nullptr, nullptr, nullptr, nullptr, prm);
#pragma unroll
          for (int f = 0; f < NVF; ++f) {
#pragma unroll
            for (int i = 0; i < BI; ++i) {
              const int p = pos[(side == 0 ? 0 : 1) * NVF * NVF + f * NVF + g];
              val[(size_t)p * (DCG_B * DCG_B) + i * DCG_B + j] += (ri2[f][i] - ri[f][i]) / delta;
            }
#pragma unroll
            for (int i = 0; i < BO; ++i) {
              const int p = pos[(side == 0 ? 2 : 3) * NVF * NVF + f * NVF + g];
              val[(size_t)p * (DCG_B * DCG_B) + i * DCG_B + j] += (ro2[f][i] - ro[f][i]) / delta;
            }
          }
        }
    }
  }
#else
  double Jii[NVF * NVF * (BI > 0 ? BI * BI : 1)], Jio[NVF * NVF * (BI * BO > 0 ? BI * BO : 1)],
      Joi[NVF * NVF * (BI * BO > 0 ? BI * BO : 1)], Joo[NVF * NVF * (BO > 0 ? BO * BO : 1)];
#pragma unroll
  for (int k = 0; k < NVF * NVF * BI * BI; ++k) Jii[k] = 0.0;
#pragma unroll
  for (int k = 0; k < NVF * NVF * BI * BO; ++k) { Jio[k] = 0.0; Joi[k] = 0.0; }
#pragma unroll
  for (int k = 0; k < NVF * NVF * BO * BO; ++k) Joo[k] = 0.0;
  facet_residual(xi, xo, fd, t, ca, ri, ro, Jii, Jio, Joi, Joo, prm);
  if (mode & 4) {
    double vi[NVF][BI > 0 ? BI : 1], vo[NVF][BO > 0 ? BO : 1];
#pragma unroll
    for (int f = 0; f < NVF; ++f) {
#pragma unroll
      for (int i = 0; i < BI; ++i) vi[f][i] = dir[(size_t)fi[f] * DCG_B + i];
#pragma unroll
      for (int i = 0; i < BO; ++i) vo[f][i] = dir[(size_t)fi[NVF + f] * DCG_B + i];
    }
#pragma unroll
    for (int f = 0; f < NVF; ++f) {
#pragma unroll
      for (int i = 0; i < BI; ++i) {
        double s = 0.0;
#pragma unroll
        for (int g = 0; g < NVF; ++g) {
#pragma unroll
          for (int j = 0; j < BI; ++j) s += Jii[((f * NVF + g) * BI + i) * BI + j] * vi[g][j];
#pragma unroll
          for (int j = 0; j < BO; ++j) s += Jio[((f * NVF + g) * BI + i) * (BO > 0 ? BO : 1) + j] * vo[g][j];
        }
        r[(size_t)fi[f] * DCG_B + i] += s;
      }
#pragma unroll
      for (int i = 0; i < BO; ++i) {
        double s = 0.0;
#pragma unroll
        for (int g = 0; g < NVF; ++g) {
#pragma unroll
          for (int j = 0; j < BI; ++j) s += Joi[((f * NVF + g) * (BO > 0 ? BO : 1) + i) * BI + j] * vi[g][j];
#pragma unroll
          for (int j = 0; j < BO; ++j)
            s += Joo[((f * NVF + g) * (BO > 0 ? BO : 1) + i) * (BO > 0 ? BO : 1) + j] * vo[g][j];
        }
        r[(size_t)fi[NVF + f] * DCG_B + i] += s;
      }
    }
  }
  if (mode & 8) {
#pragma unroll
    for (int f = 0; f < NVF; ++f) {
      double* d0 = val + (size_t)fi[f] * (DCG_B * DCG_B);
#pragma unroll
      for (int i = 0; i < BI; ++i)
#pragma unroll
        for (int j = 0; j < BI; ++j) d0[i * DCG_B + j] += Jii[((f * NVF + f) * BI + i) * BI + j];
      if (BO > 0) {
        double* d3 = val + (size_t)fi[NVF + f] * (DCG_B * DCG_B);
#pragma unroll
        for (int i = 0; i < BO; ++i)
#pragma unroll
          for (int j = 0; j < BO; ++j)
            d3[i * DCG_B + j] += Joo[((f * NVF + f) * (BO > 0 ? BO : 1) + i) * (BO > 0 ? BO : 1) + j];
      }
    }
  }
  if (want_j) {
    const int* pos = fi + 2 * NVF;
#pragma unroll
    for (int f = 0; f < NVF; ++f)
#pragma unroll
      for (int g = 0; g < NVF; ++g) {
        double* b0 = val + (size_t)pos[0 * NVF * NVF + f * NVF + g] * (DCG_B * DCG_B);
#pragma unroll
        for (int i = 0; i < BI; ++i)
#pragma unroll
          for (int j = 0; j < BI; ++j) b0[i * DCG_B + j] += Jii[((f * NVF + g) * BI + i) * BI + j];
        if (BO > 0) {
          double* b1 = val + (size_t)pos[1 * NVF * NVF + f * NVF + g] * (DCG_B * DCG_B);
          double* b2 = val + (size_t)pos[2 * NVF * NVF + f * NVF + g] * (DCG_B * DCG_B);
          double* b3 = val + (size_t)pos[3 * NVF * NVF + f * NVF + g] * (DCG_B * DCG_B);
#pragma unroll
          for (int i = 0; i < BI; ++i)
#pragma unroll
            for (int j = 0; j < BO; ++j) b1[i * DCG_B + j] += Jio[((f * NVF + g) * BI + i) * (BO > 0 ? BO : 1) + j];
#pragma unroll
          for (int i = 0; i < BO; ++i)
#pragma unroll
            for (int j = 0; j < BI; ++j) b2[i * DCG_B + j] += Joi[((f * NVF + g) * (BO > 0 ? BO : 1) + i) * BI + j];
#pragma unroll
          for (int i = 0; i < BO; ++i)
#pragma unroll
            for (int j = 0; j < BO; ++j) b3[i * DCG_B + j] += Joo[((f * NVF + g) * (BO > 0 ? BO : 1) + i) * (BO > 0 ? BO : 1) + j];
        }
      }
  }
#endif
  if (mode & 1) {
#pragma unroll
    for (int f = 0; f < NVF; ++f) {
#pragma unroll
      for (int i = 0; i < BI; ++i) r[(size_t)fi[f] * DCG_B + i] += ri[f][i];
#pragma unroll
      for (int i = 0; i < BO; ++i) r[(size_t)fi[NVF + f] * DCG_B + i] += ro[f][i];
    }
  }
}
}  // namespace %F%

extern "C" __global__ void __launch_bounds__(64)
%KERNEL%(const int* __restrict__ fint, const double* __restrict__ fdbl, int first, int count,
         const double* __restrict__ x, double t, double ca, double* __restrict__ r, double* __restrict__ val,
         int mode, const double* __restrict__ prm, const double* __restrict__ dir) {
  %F%::facet_body<double>(fint, fdbl, first, count, x, t, ca, r, val, mode, prm, dir);
}
// the FP32 apply (mode 4) of the refinement loop
extern "C" __global__ void __launch_bounds__(64)
%KERNEL%_f(const int* __restrict__ fint, const double* __restrict__ fdbl, int first, int count,
           const double* __restrict__ x, double t, double ca, float* __restrict__ r, double* __restrict__ val,
           int mode, const double* __restrict__ prm, const float* __restrict__ dir) {
  %F%::facet_body<float>(fint, fdbl, first, count, x, t, ca, r, val, mode, prm, dir);
}
)";

void replace_all(std::string& s, const std::string& from, const std::string& to) {
  size_t p = 0;
  while ((p = s.find(from, p)) != std::string::npos) {
    s.replace(p, from.size(), to);
    p += to.size();
  }
}

// variable names -> C expressions for one compartment's volume expressions
// run-time parameters follow the eight geometry words of prm
void param_vars(const ModelSpec& m, std::map<std::string, std::string>& v) {
  for (size_t k = 0; k < m.params.size(); ++k)
    v["$" + m.params[k]] = "prm[" + std::to_string(kGeoWords + k) + "]";
}

std::map<std::string, std::string> volume_vars(const ModelSpec& m, int comp, double evol, bool geo, bool uns = false) {
  std::map<std::string, std::string> v;
  param_vars(m, v);
  v["time"] = "t";
  v["entity_volume"] = geo ? "(" + D(evol) + " * prm[6])" : D(evol);
  if (uns) v["entity_volume"] = "(" + D(evol) + " * eg[UG_DET])";
  v["in_volume"] = "1.0";
  v["in_boundary"] = "0.0";
  v["in_skeleton"] = "0.0";
  v["integration_factor"] = "0.0";
  const char* pos[3] = {"px", "py", "pz"};
  for (int k = 0; k < 3; ++k) v[std::string("position_") + kAxis[k]] = k < m.dim ? pos[k] : "0.0";
  for (size_t s = 0; s < m.species.size(); ++s) {
    const auto& sp = m.species[s];
    bool own = sp.comp == comp;
    v[sp.name] = own ? "u[" + std::to_string(sp.local) + "]" : "0.0";
    for (int k = 0; k < 3; ++k)
      v["grad_" + sp.name + "_" + kAxis[k]] =
          (own && k < m.dim) ? "g[" + std::to_string(sp.local) + "][" + std::to_string(k) + "]" : "0.0";
  }
  return v;
}

// unit-spacing copy of a grid: the element tables of run-time geometry
Grid unit_cell(const Grid& g) {
  Grid u = g;
  u.h = {1.0, 1.0, 1.0};
  return u;
}

// with_kernel = false: only the compartment's tables and point function
// (what the tiled kernel builds on); the coloured kernel dominates hipRTC
// time and is left out when the tiled kernel serves the compartment.
std::string gen_volume(const ModelSpec& m, const Grid& g, int comp, const std::string& ns, const std::string& kname,
                       bool numerical, bool geo, bool with_kernel) {
  const auto T = element_tables(geo ? unit_cell(g) : g);
  const auto& species = m.comp_species[comp];
  const int bc = static_cast<int>(species.size());
  const int d = m.dim;
  auto vars = volume_vars(m, comp, T.evol, geo, g.unstructured);
  // numeric diffusion is folded into tables only when the spacing is known
  // (structured cells of a fixed size)
  auto folded = [&](const TensorTerm& t) { return !geo && !g.unstructured && tensor_const(t); };
  auto E = [&](const Ex& e) { return expr::emit(e, vars); };

  bool has_gc = false, has_dc = false;
  for (int s : species) {
    const auto& sp = m.species[s];
    for (int k = 0; k < d; ++k) has_gc = has_gc || sp.velocity[k];
    for (const auto& dt : sp.diffusion) {
      if (!dt.jac.empty()) has_gc = true;
      if (!folded(dt.D)) has_dc = true;
    }
  }
  std::ostringstream o;
  o << "namespace " << ns << " {\n";
  o << "constexpr int BC = " << bc << ", NV = " << T.nv << ", NQ = " << T.nq << ", NCORNER = " << g.ncorner
    << ", NCLASS = " << g.nclass << ", DIM = " << d << ";\n";
  // per-element geometry words of unstructured cells: basis gradients, points, |det J|
  o << "constexpr int UG_POS = NV * DIM, UG_DET = UG_POS + NQ * 3, UG_WORDS = UG_DET + 1;\n";
  o << "constexpr double FD_EPS = " << D(m.fd_epsilon) << ";\n";
  std::vector<int> corner;
  for (int p = 0; p < g.nclass; ++p)
    for (int a = 0; a < T.nv; ++a) corner.push_back(g.corner[p][a]);
  o << "__device__ constexpr int CORNER[NCLASS][NV] = " << int_init(corner) << ";\n";
  o << "__device__ constexpr int OFF[NCLASS][NV][NV] = " << int_init(T.off) << ";\n";
  o << "__device__ constexpr double PHI[NQ][NV] = " << array_init(T.phi) << ";\n";
  o << "__device__ constexpr double GPHI[NCLASS][NQ][NV][DIM] = " << array_init(T.gphi) << ";\n";
  o << "__device__ constexpr double FAC[NQ] = " << array_init(T.fac) << ";\n";
  o << "__device__ constexpr double XQ[NCLASS][NQ][3] = " << array_init(T.xq) << ";\n";
  // constant diffusion: sum_q FAC grad phi_a . D grad phi_c folded per species pair
  {
    // KD[p][a][c][i][j]
    std::vector<double> kd(static_cast<size_t>(g.nclass) * T.nv * T.nv * std::max(1, bc * bc), 0.0);
    bool any = false;
    for (int i = 0; i < bc; ++i) {
      const auto& sp = m.species[species[i]];
      for (const auto& dt : sp.diffusion) {
        if (!folded(dt.D)) continue;
        int j = m.species[dt.wrt].local;
        double Dm[3][3] = {{0}};
        for (int k = 0; k < d; ++k)
          for (int l = 0; l < d; ++l) {
            double v = 0;
            if (dt.D.tensor) expr::is_const(dt.D.d[k][l], &v);
            else if (k == l) expr::is_const(dt.D.d[0][0], &v);
            Dm[k][l] = v;
          }
        for (int p = 0; p < g.nclass; ++p)
          for (int a = 0; a < T.nv; ++a)
            for (int c = 0; c < T.nv; ++c) {
              double s = 0;
              for (int q = 0; q < T.nq; ++q)
                for (int k = 0; k < d; ++k)
                  for (int l = 0; l < d; ++l) {
                    double ga = T.gphi[((static_cast<size_t>(p) * T.nq + q) * T.nv + a) * d + k];
                    double gc = T.gphi[((static_cast<size_t>(p) * T.nq + q) * T.nv + c) * d + l];
                    s += T.fac[q] * ga * Dm[k][l] * gc;
                  }
              kd[((static_cast<size_t>(p) * T.nv + a) * T.nv + c) * bc * bc + i * bc + j] += s;
              any = any || s != 0.0;
            }
      }
    }
    if (any) {
      o << "__device__ constexpr double KD[NCLASS][NV][NV][BC][BC] = " << array_init(kd) << ";\n";
      o << "template <int P> __device__ __forceinline__ double const_diffusion(int a, int c, int i, int j) { return KD[P][a][c][i][j]; }\n";
      o << "__device__ __forceinline__ double kd_rt(int p, int a, int c, int i, int j) { return KD[p][a][c][i][j]; }\n";
    } else {
      o << "template <int P> __device__ __forceinline__ double const_diffusion(int, int, int, int) { return 0.0; }\n";
      o << "__device__ __forceinline__ double kd_rt(int, int, int, int, int) { return 0.0; }\n";
    }
  }
  // quadrature point kernel
  o << "__device__ __forceinline__ void qp(int q, const double (&u)[BC], const double (&g)[BC][3], double px, double py, "
       "double pz, double t, double cm, double ca, double (&rs)[BC], double (&rf)[BC][DIM], "
    << (with_kernel ? "double (*mc)[BC], " : "double* mc, ")
    << "double (*gc)[BC][DIM], double (*dc)[BC][DIM][DIM], const double* __restrict__ prm, "
       "const double* __restrict__ eg) {\n";
  o << "  (void)px; (void)py; (void)pz; (void)t; (void)g; (void)gc; (void)dc; (void)prm; (void)eg;\n";
  o << "  const double fac = FACQ(q);\n";
  o << "#pragma unroll\n  for (int i = 0; i < BC; ++i) { rs[i] = 0.0;\n#pragma unroll\n for (int k = 0; k < DIM; ++k) rf[i][k] = 0.0; }\n";
  // tiled compartments keep only the structural nonzeros of mc, in the
  // order of mc_pattern (the tiled kernel's MCI / MCJ tables)
  const auto mcnz = mc_pattern(m, comp);
  auto MC = [&](int i, int j) {
    if (with_kernel) return "mc[" + std::to_string(i) + "][" + std::to_string(j) + "]";
    for (size_t n = 0; n < mcnz.size(); ++n)
      if (mcnz[n].first == i && mcnz[n].second == j) return "mc[" + std::to_string(n) + "]";
    throw Error("internal: coefficient outside the point pattern");
  };
  if (with_kernel)
    o << "  if (mc) {\n#pragma unroll\n  for (int i = 0; i < BC; ++i)\n#pragma unroll\n for (int j = 0; j < BC; ++j) mc[i][j] = 0.0;}\n";
  else if (!mcnz.empty())
    o << "  if (mc) {\n#pragma unroll\n  for (int n = 0; n < " << mcnz.size() << "; ++n) mc[n] = 0.0;}\n";
  if (has_gc)
    o << "  if (gc) {\n#pragma unroll\n for (int i = 0; i < BC; ++i)\n#pragma unroll\n for (int j = 0; j < BC; ++j)\n#pragma unroll\n for (int k = 0; k < DIM; ++k) gc[i][j][k] = 0.0;}\n";
  if (has_dc)
    o << "  if (dc) {\n#pragma unroll\n for (int i = 0; i < BC; ++i)\n#pragma unroll\n for (int j = 0; j < BC; ++j)\n#pragma unroll\n for (int k = 0; k < DIM; ++k)\n#pragma unroll\n for (int l = 0; l < DIM; ++l) dc[i][j][k][l] = 0.0;}\n";
  for (int i = 0; i < bc; ++i) {
    const auto& sp = m.species[species[i]];
    const std::string I = std::to_string(i);
    o << "  { // species " << sp.name << "\n";
    if (sp.storage) {
      o << "    const double S = " << E(sp.storage) << ";\n";
      o << "    rs[" << I << "] += fac * cm * S * u[" << I << "];\n";
      o << "    if (mc) { " << MC(i, i) << " += fac * cm * S;\n";
      for (const auto& pj : sp.storage_jac) {
        if (m.species[pj.wrt].comp != comp) continue;
        o << "      " << MC(i, m.species[pj.wrt].local) << " += fac * cm * (" << E(pj.e) << ") * u[" << I << "];\n";
      }
      o << "    }\n";
    }
    if (sp.reaction) {
      o << "    rs[" << I << "] -= fac * ca * (" << E(sp.reaction) << ");\n";
      o << "    if (mc) {\n";
      for (const auto& pj : sp.reaction_jac) {
        if (m.species[pj.wrt].comp != comp) continue;
        o << "      " << MC(i, m.species[pj.wrt].local) << " -= fac * ca * (" << E(pj.e) << ");\n";
      }
      o << "    }\n";
    }
    bool hv = false;
    for (int k = 0; k < d; ++k) hv = hv || sp.velocity[k];
    if (hv) {
      for (int k = 0; k < d; ++k) {
        if (!sp.velocity[k]) continue;
        o << "    { const double vk = " << E(sp.velocity[k]) << ";\n";
        o << "      rf[" << I << "][" << k << "] -= fac * ca * vk * u[" << I << "];\n";
        o << "      if (gc) gc[" << I << "][" << I << "][" << k << "] -= fac * ca * vk; }\n";
      }
      for (const auto& vj : sp.velocity_jac) {
        int j = m.species[vj.first].local;
        for (int k = 0; k < d; ++k)
          if (vj.second[k])
            o << "    if (gc) gc[" << I << "][" << j << "][" << k << "] -= fac * ca * (" << E(vj.second[k]) << ") * u[" << I
              << "];\n";
      }
    }
    for (const auto& dt : sp.diffusion) {
      int j = m.species[dt.wrt].local;
      const std::string J = std::to_string(j);
      // flux D grad u_j for the residual
      o << "    {\n";
      for (int k = 0; k < d; ++k)
        for (int l = 0; l < d; ++l) {
          Ex e = dt.D.tensor ? dt.D.d[k][l] : (k == l ? dt.D.d[0][0] : nullptr);
          if (!e) continue;
          o << "      { const double Dv = " << E(e) << ";\n";
          o << "        rf[" << I << "][" << k << "] += fac * ca * Dv * g[" << J << "][" << l << "];\n";
          if (!folded(dt.D)) o << "        if (dc) dc[" << I << "][" << J << "][" << k << "][" << l << "] += fac * ca * Dv;\n";
          o << "      }\n";
        }
      for (const auto& dj : dt.jac) {
        int v = m.species[dj.first].local;
        for (int k = 0; k < d; ++k)
          for (int l = 0; l < d; ++l) {
            Ex e = dj.second.tensor ? dj.second.d[k][l] : (k == l ? dj.second.d[0][0] : nullptr);
            if (!e) continue;
            o << "      if (gc) gc[" << I << "][" << v << "][" << k << "] += fac * ca * (" << E(e) << ") * g[" << J << "][" << l
              << "];\n";
          }
      }
      o << "    }\n";
    }
    o << "  }\n";
  }
  o << "}\n}  // namespace " << ns << "\n";
  std::string body = kVolume;
  replace_all(body, "%C%", ns);
  replace_all(body, "%KERNEL%", kname);
  std::ostringstream disp;
  for (int p = 0; p < g.nclass; ++p)
    disp << "  if ((msk >> " << p << ") & 1u) " << ns << "::element<" << p
         << ">(xc, sl, cx, cy, cz, t, cm, ca, r, val, own_start, own_mask, mode, prm, eg);\n";
  replace_all(body, "%DISPATCH%", disp.str());
  std::ostringstream full;
  full << "#define NUMERICAL_JAC " << (numerical ? 1 : 0) << "\n#define HAS_GC " << (has_gc ? 1 : 0) << "\n#define HAS_DC "
       << (has_dc ? 1 : 0) << "\n";
  full << o.str() << (with_kernel ? body : std::string()) << "#undef NUMERICAL_JAC\n#undef HAS_GC\n#undef HAS_DC\n";
  return full.str();
}

std::string gen_facet(const ModelSpec& m, const Grid& g, int cin, int cout, bool boundary, const std::string& ns,
                      const std::string& kname, bool numerical, FacetKernel& fk) {
  Rule r = facet_rule(g.dim, g.cube);
  const int nqf = static_cast<int>(r.w.size());
  const int nvf = g.cube ? (1 << (g.dim - 1)) : g.dim;
  const int bi = static_cast<int>(m.comp_species[cin].size());
  const int bo = boundary ? 0 : static_cast<int>(m.comp_species[cout].size());
  fk.nvf = nvf;
  fk.nqf = nqf;
  fk.fint = 2 * nvf + 4 * nvf * nvf;
  fk.fdbl = 4 * nqf + 4;
  // facet basis values (same vertex-order convention as the layout)
  std::vector<double> phif;
  for (int q = 0; q < nqf; ++q)
    for (int f = 0; f < nvf; ++f) {
      double e0 = r.x[q][0], e1 = r.x[q][1], phi;
      if (g.dim == 1) phi = 1.0;
      else if (g.cube) {
        phi = (f & 1) ? e0 : 1.0 - e0;
        if (g.dim == 3) phi *= ((f >> 1) & 1) ? e1 : 1.0 - e1;
      } else {
        phi = f == 0 ? 1.0 - e0 - (g.dim == 3 ? e1 : 0.0) : (f == 1 ? e0 : e1);
      }
      phif.push_back(phi);
    }
  auto vars_for = [&](bool inside) {
    std::map<std::string, std::string> v;
    param_vars(m, v);
    v["time"] = "t";
    v["entity_volume"] = "fvol";
    v["in_volume"] = "0.0";
    v["in_boundary"] = boundary ? "1.0" : "0.0";
    v["in_skeleton"] = boundary ? "0.0" : "1.0";
    v["integration_factor"] = "0.0";
    const char* pos[3] = {"px", "py", "pz"};
    const char* nrm[3] = {"nx", "ny", "nz"};
    for (int k = 0; k < 3; ++k) {
      v[std::string("position_") + kAxis[k]] = k < m.dim ? pos[k] : "0.0";
      v[std::string("normal_") + kAxis[k]] = k < m.dim ? (inside ? std::string(nrm[k]) : "(-" + std::string(nrm[k]) + ")") : "0.0";
    }
    for (size_t s = 0; s < m.species.size(); ++s) {
      const auto& sp = m.species[s];
      std::string val = "0.0";
      if (sp.comp == cin) val = "ui[" + std::to_string(sp.local) + "]";
      else if (!boundary && sp.comp == cout) val = "uo[" + std::to_string(sp.local) + "]";
      v[sp.name] = val;
      // gradients on facets are not provided by this build
    }
    return v;
  };
  std::ostringstream qb;
  auto side = [&](bool inside) {
    int own = inside ? cin : cout;
    int target = boundary ? cin : (inside ? cout : cin);
    auto vars = vars_for(inside);
    const std::string R = inside ? "ri" : "ro";
    const int bown = inside ? bi : bo;
    const std::string Jsame = inside ? "Jii" : "Joo";
    const std::string Jother = inside ? "Jio" : "Joi";
    const int bother = inside ? bo : bi;
    for (int s : m.comp_species[own]) {
      const auto& sp = m.species[s];
      for (const auto& of : sp.outflow) {
        if (of.to_comp != target) continue;
        std::vector<std::string> used;
        expr::collect_vars(of.e, used);
        for (auto& p : of.jac) expr::collect_vars(p.e, used);
        for (auto& u : used)
          if (u.rfind("grad_", 0) == 0) throw NotImplemented("species gradients in outflow expressions");
        const int i = sp.local;
        qb << "    { const double T = " << expr::emit(of.e, vars) << ";\n";
        qb << "#pragma unroll\n      for (int f = 0; f < NVF; ++f) " << R << "[f][" << i << "] += w * T * PHIF[q][f];\n";
        qb << "      if (Jii) {\n";
        for (const auto& pj : of.jac) {
          const auto& wsp = m.species[pj.wrt];
          std::string Jn;
          int j = wsp.local, bcol;
          if (wsp.comp == own) {
            Jn = Jsame;
            bcol = bown;
          } else if (!boundary && wsp.comp == (inside ? cout : cin)) {
            Jn = Jother;
            bcol = bother;
          } else {
            continue;
          }
          qb << "        { const double dT = w * (" << expr::emit(pj.e, vars) << ");\n";
          qb << "#pragma unroll\n          for (int f = 0; f < NVF; ++f)\n#pragma unroll\n            for (int gg = 0; gg < NVF; ++gg) "
             << Jn << "[((f * NVF + gg) * " << std::max(bown, 1) << " + " << i << ") * " << std::max(bcol, 1) << " + " << j
             << "] += dT * PHIF[q][f] * PHIF[q][gg]; }\n";
        }
        qb << "      }\n    }\n";
      }
    }
    if (!inside) return;
  };
  qb << "    {\n";
  side(true);
  qb << "    }\n";
  if (!boundary) {
    qb << "    {\n";
    side(false);
    qb << "    }\n";
  }
  std::ostringstream o;
  o << "namespace " << ns << " {\n";
  o << "constexpr int NVF = " << nvf << ", NQF = " << nqf << ", BI = " << bi << ", BO = " << bo << ", FINT = " << fk.fint
    << ", FDBL = " << fk.fdbl << ";\n";
  o << "constexpr double FD_EPS = " << D(m.fd_epsilon) << ";\n";
  o << "__device__ constexpr double PHIF[NQF][NVF] = " << array_init(phif) << ";\n";
  o << "}\n";
  std::string body = kFacet;
  replace_all(body, "%F%", ns);
  replace_all(body, "%KERNEL%", kname);
  replace_all(body, "%QPBODY%", qb.str());
  std::ostringstream full;
  full << "#define NUMERICAL_JAC " << (numerical ? 1 : 0) << "\n" << o.str() << body << "#undef NUMERICAL_JAC\n";
  return full.str();
}


// Tiled write-once volume kernel.  One workgroup owns a TX x TY x TZ box of
// vertices; it stages the state of the surrounding vertex box and the class
// masks of the surrounding cells in LDS, computes per (cell, class, point)
// Jacobian coefficients (or per-cell element residuals) once, and then every
// thread gathers one BCSR block (or one residual entry) of its rows from the
// cells that touch it and stores it exactly once.  No zero fill, no colours,
// no atomics; the summation order is fixed.
const char* kTiled = R"(
namespace %C% {
constexpr int XB = TX + 2, YB = DIM > 1 ? TY + 2 : 1, ZB = DIM > 2 ? TZ + 2 : 1;
constexpr int CXB = TX + 1, CYB = DIM > 1 ? TY + 1 : 1, CZB = DIM > 2 ? TZ + 1 : 1;
constexpr int NVB = XB * YB * ZB, NCB = CXB * CYB * CZB, NROW = TX * TY * TZ;
constexpr int NST = DIM == 1 ? 3 : DIM == 2 ? 9 : 27;
constexpr int OY = DIM > 1 ? 1 : 0, OZ = DIM > 2 ? 1 : 0;
constexpr int MC_WORDS = NCB * NCLASS * NQ * (NMC > 0 ? NMC : 1);
// the cells a tile owns (their lowest corner is one of its vertices): each
// cell's stored point coefficients live once, in its owner's block
constexpr int NOWN = TX * (DIM > 1 ? TY : 1) * (DIM > 2 ? TZ : 1);
constexpr int OWN_WORDS = NCLASS * NQ * (NMC > 0 ? NMC : 1) * NOWN;
// residual corner sums: component-major for box cells with blocks of up to
// 4 species, else per cell with an odd stride (RL_CELL doubles)
constexpr int RL_CELL = NCORNER * BC + 1;
constexpr int RL_WORDS = NCB * RL_CELL;
// the Jacobian paths also pass row sums between the two lanes of a row
constexpr int XCH_WORDS = NROW * ((NZB > BC ? NZB : BC) > 0 ? (NZB > BC ? NZB : BC) : 1);
constexpr int JW_WORDS = MC_WORDS > XCH_WORDS ? MC_WORDS : XCH_WORDS;
constexpr int UNION_WORDS = JW_WORDS > RL_WORDS ? JW_WORDS : RL_WORDS;

// element residual of class p (runtime) from corner values
__device__ __forceinline__ void residual_rt(int p, const double (&xl)[NV][BC], double cx, double cy, double cz,
                                            double t, double cm, double ca, double (&rl)[NV][BC],
                                            const double* __restrict__ prm) {
  for (int a = 0; a < NV; ++a)
    for (int i = 0; i < BC; ++i) rl[a][i] = 0.0;
  for (int q = 0; q < NQ; ++q) {
    double u[BC], g[BC][3];
    for (int i = 0; i < BC; ++i) {
      double s = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
      for (int a = 0; a < NV; ++a) {
        s += PHI[q][a] * xl[a][i];
        g0 += GP(p, q, a, 0) * xl[a][i];
        if (DIM > 1) g1 += GP(p, q, a, DIM > 1 ? 1 : 0) * xl[a][i];
        if (DIM > 2) g2 += GP(p, q, a, DIM > 2 ? 2 : 0) * xl[a][i];
      }
      u[i] = s;
      g[i][0] = g0;
      g[i][1] = g1;
      g[i][2] = g2;
    }
    const double px = cx + XQP(p, q, 0), py = cy + XQP(p, q, 1), pz = cz + XQP(p, q, 2);
    double rs[BC], rf[BC][DIM];
    qp(q, u, g, px, py, pz, t, cm, ca, rs, rf, nullptr, nullptr, nullptr, prm, nullptr);
    for (int a = 0; a < NV; ++a)
      for (int i = 0; i < BC; ++i) {
        double s = rs[i] * PHI[q][a];
        for (int k = 0; k < DIM; ++k) s += rf[i][k] * GP(p, q, a, k);
        rl[a][i] += s;
      }
  }
}

// element residual of class p accumulated into dst[corner * BC + i], the
// corner values read from the staged box xs at bidx[a] (for large blocks,
// which would otherwise hold two NV x BC arrays in registers)
__device__ __forceinline__ void residual_lds(int p, const double* __restrict__ xs, const int (&bidx)[NV], double cx,
                                             double cy, double cz, double t, double cm, double ca,
                                             double* __restrict__ dst, const double* __restrict__ prm) {
  // one point at a time, the corner values re-read from LDS for each point
  // (the empty asm keeps the compiler from holding them in registers)
#pragma unroll 1
  for (int q = 0; q < NQ; ++q) {
    __asm__ volatile("" ::: "memory");
    double u[BC], g[BC][3];
    for (int i = 0; i < BC; ++i) {
      u[i] = 0.0;
      g[i][0] = g[i][1] = g[i][2] = 0.0;
    }
    for (int a = 0; a < NV; ++a) {
      const double pa = PHI[q][a], ga0 = GP(p, q, a, 0);
      const double ga1 = DIM > 1 ? GP(p, q, a, DIM > 1 ? 1 : 0) : 0.0, ga2 = DIM > 2 ? GP(p, q, a, DIM > 2 ? 2 : 0) : 0.0;
      for (int i = 0; i < BC; ++i) {
        const double xv = xs[bidx[a] * BC + i];
        u[i] += pa * xv;
        g[i][0] += ga0 * xv;
        if (DIM > 1) g[i][1] += ga1 * xv;
        if (DIM > 2) g[i][2] += ga2 * xv;
      }
    }
    const double px = cx + XQP(p, q, 0), py = cy + XQP(p, q, 1), pz = cz + XQP(p, q, 2);
    double rs[BC], rf[BC][DIM];
    qp(q, u, g, px, py, pz, t, cm, ca, rs, rf, nullptr, nullptr, nullptr, prm, nullptr);
    for (int a = 0; a < NV; ++a) {
      double* d = dst + CORNER[p][a] * BC;
      for (int i = 0; i < BC; ++i) {
        double v = rs[i] * PHI[q][a];
        for (int k = 0; k < DIM; ++k) v += rf[i][k] * GP(p, q, a, k);
        d[i] += v;
      }
    }
  }
}

// block rows i0 .. i0 + AI - 1 of the Jacobian block that couples the row
// vertex (li, lj, lk) of the tile to its neighbour at offset (dx, dy, dz):
// the staged point coefficients of the cells around the row vertex that
// contain both vertices, weighted by the reference tables, plus diffusion
template <int AI>
__device__ __forceinline__ void block_rows(int dx, int dy, int dz, int li, int lj, int lk, int i0,
                                           const unsigned char* __restrict__ cms, const double* __restrict__ work,
                                           const double* __restrict__ dv, double ca, double (&acc)[AI][BC]) {
#pragma unroll
  for (int i = 0; i < AI; ++i)
#pragma unroll
    for (int j = 0; j < BC; ++j) acc[i][j] = 0.0;
  if (i0 >= BC) return;
  for (int ez = 0; ez < (DIM > 2 ? 2 : 1); ++ez) {
    if (DIM > 2 && ((dz == 1 && ez == 0) || (dz == -1 && ez == 1))) continue;
    for (int ey = 0; ey < (DIM > 1 ? 2 : 1); ++ey) {
      if (DIM > 1 && ((dy == 1 && ey == 0) || (dy == -1 && ey == 1))) continue;
      for (int ex = 0; ex < 2; ++ex) {
        if ((dx == 1 && ex == 0) || (dx == -1 && ex == 1)) continue;
        const int cl = (li + ex) + CXB * ((lj + ey) + CYB * (lk + ez));
        const unsigned m = cms[cl];
        if (!m) continue;
        // corner bits of the row vertex and of its neighbour in this cell
        const int cv = (1 - ex) | (DIM > 1 ? (1 - ey) << 1 : 0) | (DIM > 2 ? (1 - ez) << 2 : 0);
        const int cw = (1 - ex + dx) | (DIM > 1 ? (1 - ey + dy) << 1 : 0) | (DIM > 2 ? (1 - ez + dz) << 2 : 0);
        for (int p = 0; p < NCLASS; ++p) {
          const int la = LOC[p][cv], lc = LOC[p][cw];
          if (la < 0 || lc < 0) continue;
          if (!((m >> p) & 1u)) continue;
          for (int q = 0; q < NQ; ++q) {
            const double w = PP[q][la][lc];
            const double* mcp = work + (size_t)((p * NQ + q) * NMC) * NCB + cl;
#pragma unroll
            for (int n = 0; n < NMC; ++n) {
              const int ri = MCI[n] - i0;
              if (ri >= 0 && ri < AI) acc[AI == 1 ? 0 : (ri < AI ? ri : 0)][MCJ[n]] += w * mcp[(size_t)n * NCB];
            }
          }
          if constexpr (CONST_DIFFUSION) {
#pragma unroll
            for (int i = 0; i < AI; ++i)
#pragma unroll
              for (int j = 0; j < BC; ++j) acc[i][j] += ca * kd_rt(p, la, lc, i0 + i, j);
          } else {
#pragma unroll
            for (int n = 0; n < NDV; ++n) {
              const int ri = DVI[n] - i0;
              if (ri >= 0 && ri < AI)
                acc[AI == 1 ? 0 : (ri < AI ? ri : 0)][DVJ[n]] += ca * dv[n] * KDU[p][la][lc][DVK[n]][DVL[n]];
            }
          }
        }
      }
    }
  }
}

// The box geometry of one tile of TX x TY x TZ row vertices: the vertex box
// around it (XB x YB x ZB) and the cells touching its rows (CXB x CYB x CZB).
// The matrix-free apply uses the assembly tile; the point apply has its own,
// chosen per model by timing the candidates.
template <int X, int Y, int Z, int THREADS = 256, int WAVES = 1>
struct ApplyTile {
  static constexpr int NT = THREADS;  // workgroup size
  static constexpr int MINW = WAVES;  // waves per SIMD the registers are bounded for
  static constexpr int TX = X, TY = DIM > 1 ? Y : 1, TZ = DIM > 2 ? Z : 1;
  static constexpr int XB = TX + 2, YB = DIM > 1 ? TY + 2 : 1, ZB = DIM > 2 ? TZ + 2 : 1;
  static constexpr int CXB = TX + 1, CYB = DIM > 1 ? TY + 1 : 1, CZB = DIM > 2 ? TZ + 1 : 1;
  static constexpr int NVB = XB * YB * ZB, NCB = CXB * CYB * CZB, NROW = TX * TY * TZ;
};
%APPLY_SHAPES%
// The rows of one tile from point coefficients coef(box cell, (p*NQ + q)*NMC +
// n) (read from the grid-wide array in HBM, or the set-up's staged LDS array)
// and the direction staged in vs over the vertex box (slots sl, class masks cms): per
// cell, the direction at its corners interpolated to each point, multiplied
// by the point's block and tested with the basis functions, plus the uniform
// diffusion's element stiffness times the corner values; the cells' corner
// sums folded onto the tile's rows.  y gets the rows (padding species rows
// are the identity, from v); with part, the workgroup's partial dots (y.w1,
// y.w2) go to part[0] and part[1].  LDS scratch: accv (zeroed by the
// caller) when SCATTER, else out.
// VP, YP: pointer types of the vectors read and of y (global-memory pointers
// from the point apply, plain ones from the matrix-free apply).
template <typename TL, typename T, typename CF, typename VP, typename YP>
__device__ __forceinline__ void cells_apply(CF coef, const T* vs, const int* sl,
                                            const unsigned char* cms, T* out, T* accv, VP __restrict__ v,
                                            YP __restrict__ y, double ca, const double (&dv)[NDV > 0 ? NDV : 1],
                                            VP __restrict__ w1, VP __restrict__ w2,
                                            double* __restrict__ part) {
  constexpr int TX = TL::TX, TY = TL::TY, XB = TL::XB, YB = TL::YB, CXB = TL::CXB, CYB = TL::CYB;
  constexpr int NCB = TL::NCB, NROW = TL::NROW;
  // corner values stay in registers for small blocks, in LDS otherwise; with
  // small blocks and one cell per thread the corner sums go straight onto
  // the vertices (SCATTER), otherwise through a per-cell corner array
  constexpr bool VREG = BC <= 4;
  constexpr bool SCATTER = VREG && NCB <= TL::NT;
  // the corner sums of box cell cl over its classes, handed to put(corner, i, value)
  auto cell_corner_sums = [&](int cl, unsigned m, auto&& put) {
    const int ci = cl % CXB, cj = (cl / CXB) % CYB, ck = cl / (CXB * CYB);
#pragma unroll
    for (int p = 0; p < NCLASS; ++p) {
      if (!((m >> p) & 1u)) continue;
      int bidx[NV];
      T vc[VREG ? NV : 1][BC];
#pragma unroll
      for (int a = 0; a < NV; ++a) {
        const int c = CORNER[p][a];
        bidx[a] = (ci + (c & 1)) + XB * ((cj + ((c >> 1) & 1)) + YB * (ck + ((c >> 2) & 1)));
        if constexpr (VREG)
#pragma unroll
          for (int i = 0; i < BC; ++i) vc[a][i] = vs[bidx[a] * BC + i];
      }
      auto vcor = [&](int a, int i) -> T { return VREG ? vc[VREG ? a : 0][i] : vs[bidx[a] * BC + i]; };
      // the point terms of one point into the corner sums: acc(a, i) += PHI[q][a] w[i]
      auto point_terms = [&](int q, auto&& add) {
        T vq[BC];
#pragma unroll
        for (int j = 0; j < BC; ++j) {
          T sum = T(0);
#pragma unroll
          for (int a = 0; a < NV; ++a) sum += T(PHI[q][a]) * vcor(a, j);
          vq[j] = sum;
        }
        const int w0 = (p * NQ + q) * NMC;
        T w[BC];
#pragma unroll
        for (int i = 0; i < BC; ++i) w[i] = T(0);
#pragma unroll
        for (int n = 0; n < NMC; ++n) w[MCI[n]] += T(coef(cl, w0 + n)) * vq[MCJ[n]];
#pragma unroll
        for (int a = 0; a < NV; ++a)
#pragma unroll
          for (int i = 0; i < BC; ++i) add(a, i, T(PHI[q][a]) * w[i]);
      };
      // the uniform diffusion: acc(a, i) += ca sum_c K[a][c] v_c for the element
      // stiffness K (the table entries are immediates; ca scales the sum)
      auto diffusion_terms = [&](auto&& add) {
#pragma unroll
        for (int a = 0; a < NV; ++a) {
          if constexpr (CONST_DIFFUSION) {
#pragma unroll
            for (int n = 0; n < NKZ; ++n) {
              T sum = T(0);
#pragma unroll
              for (int c = 0; c < NV; ++c) sum += T(kd_rt(p, a, c, KZI[n], KZJ[n])) * vcor(c, KZJ[n]);
              add(a, KZI[n], T(ca) * sum);
            }
          } else {
#pragma unroll
            for (int n = 0; n < NDV; ++n) {
              T sum = T(0);
#pragma unroll
              for (int c = 0; c < NV; ++c) sum += T(KDU[p][a][c][DVK[n]][DVL[n]]) * vcor(c, DVJ[n]);
              add(a, DVI[n], T(ca * dv[n]) * sum);
            }
          }
        }
      };
      if constexpr (VREG) {
        // small blocks: the class's corner sums in registers
        T o[NV][BC];
#pragma unroll
        for (int a = 0; a < NV; ++a)
#pragma unroll
          for (int i = 0; i < BC; ++i) o[a][i] = T(0);
        auto add = [&](int a, int i, T v2) { o[a][i] += v2; };
#pragma unroll
        for (int q = 0; q < NQ; ++q) point_terms(q, add);
        diffusion_terms(add);
#pragma unroll
        for (int a = 0; a < NV; ++a)
#pragma unroll
          for (int i = 0; i < BC; ++i) put(CORNER[p][a], i, o[a][i]);
      } else {
        // large blocks: straight into the sums, point by point
        auto add = [&](int a, int i, T v2) { put(CORNER[p][a], i, v2); };
#pragma unroll 1
        for (int q = 0; q < NQ; ++q) point_terms(q, add);
        diffusion_terms(add);
      }
    }
  };
  if constexpr (SCATTER) {
    // one cell per thread; its corner sums are added to the vertices corner
    // by corner: for one corner the cells' vertices are distinct, so no two
    // threads add to the same vertex between two barriers, and every vertex
    // sums its cells in a fixed order.  When a box row of cells lies within
    // one wave (CXB divides 64), the x neighbour's corner sums arrive by lane
    // shuffle instead: lane ci holds vertex x = ci's sums of cells ci-1 and
    // ci (own first), which halves the rounds; when a whole xy layer lies
    // within one wave the y neighbour's arrive the same way (CXB lanes away)
    // and the rounds halve again.  Only the tile's own rows are read back,
    // so the halo vertices (x = 0 and x = CXB, likewise in y) need no sums.
    constexpr bool CMB_X = DIM > 1 && 64 % CXB == 0;
    constexpr bool CMB_Y = CMB_X && 64 % (CXB * CYB) == 0;
    constexpr int NFOLD = CMB_Y ? 4 : CMB_X ? 2 : 1;  // corners folded into one round
    constexpr int NR = NCORNER / NFOLD;             // rounds
    const int cl = threadIdx.x;
    const bool in = cl < NCB;
    const bool has = in && cms[in ? cl : 0];
    T oc[NCORNER][BC];
#pragma unroll
    for (int c = 0; c < NCORNER; ++c)
#pragma unroll
      for (int i = 0; i < BC; ++i) oc[c][i] = T(0);
    if (has) cell_corner_sums(cl, cms[cl], [&](int c, int i, T v2) { oc[c][i] += v2; });
    const int ci = cl % CXB, cj = (cl / CXB) % CYB, ck = cl / (CXB * CYB);
    // rs[r]: the sums of round r's vertex; round r takes corner r * NFOLD
    T rs[NR][BC];
    if constexpr (CMB_X) {
      T fx[NCORNER / 2][BC];
#pragma unroll
      for (int h = 0; h < NCORNER / 2; ++h)
#pragma unroll
        for (int i = 0; i < BC; ++i) fx[h][i] = oc[2 * h][i] + __shfl_up(oc[2 * h + 1][i], 1, 64);
      if constexpr (CMB_Y) {
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
          for (int i = 0; i < BC; ++i) rs[r][i] = fx[2 * r][i] + __shfl_up(fx[2 * r + 1][i], CXB, 64);
      } else {
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
          for (int i = 0; i < BC; ++i) rs[r][i] = fx[r][i];
      }
    } else {
#pragma unroll
      for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int i = 0; i < BC; ++i) rs[r][i] = oc[r][i];
    }
    const bool add = CMB_Y ? in && ci > 0 && cj > 0 : CMB_X ? in && ci > 0 : has;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (r) __syncthreads();
      if (add) {
        const int c = r * NFOLD;
        const int b = (ci + (c & 1)) + XB * ((cj + ((c >> 1) & 1)) + YB * (ck + ((c >> 2) & 1)));
#pragma unroll
        for (int i = 0; i < BC; ++i) accv[b * BC + i] += rs[r][i];
      }
    }
  } else {
    for (int cl = threadIdx.x; cl < NCB; cl += blockDim.x) {
      T* dst = out + (size_t)cl * (NCORNER * BC);
      for (int k = 0; k < NCORNER * BC; ++k) dst[k] = T(0);
      const unsigned m = cms[cl];
      if (!m) continue;
      cell_corner_sums(cl, m, [&](int c, int i, T v2) { dst[c * BC + i] += v2; });
    }
  }
  __syncthreads();
  double d1 = 0.0, d2 = 0.0;
  for (int it = threadIdx.x; it < NROW * DCG_B; it += blockDim.x) {
    const int lr = it / DCG_B, i = it % DCG_B;
    const int li = lr % TX, lj = DIM > 1 ? (lr / TX) % TY : 0, lk = DIM > 2 ? lr / (TX * TY) : 0;
    const int vb = (li + 1) + XB * ((lj + OY) + YB * (lk + OZ));
    const int s = sl[vb];
    if (s < 0) continue;
    T acc = T(0);
    if (i < BC && SCATTER) {
      acc = accv[(SCATTER ? vb : 0) * BC + i];
    } else if (i < BC) {
      for (int ez = 0; ez < (DIM > 2 ? 2 : 1); ++ez)
        for (int ey = 0; ey < (DIM > 1 ? 2 : 1); ++ey)
          for (int ex = 0; ex < 2; ++ex) {
            const int cl = (li + ex) + CXB * ((lj + ey) + CYB * (lk + ez));
            if (!cms[cl]) continue;
            const int cv = (1 - ex) | (DIM > 1 ? (1 - ey) << 1 : 0) | (DIM > 2 ? (1 - ez) << 2 : 0);
            acc += out[(SCATTER ? 0 : (size_t)cl * (NCORNER * BC) + cv * BC) + i];
          }
    } else {
      acc = v[(size_t)s * DCG_B + i];  // padding species: identity rows
    }
    y[(size_t)s * DCG_B + i] = acc;
    if (part) {
      // a dot operand may be y itself (t.t of BiCGSTAB): then it is acc
      d1 += (double)acc * (double)(w1 == y ? acc : w1[(size_t)s * DCG_B + i]);
      if (w2) d2 += (double)acc * (double)(w2 == y ? acc : w2[(size_t)s * DCG_B + i]);
    }
  }
  if (part) {
    // workgroup sums of the two partial dots (fixed order)
    __shared__ double red[32];  // two sums per wave, up to 16 waves
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      d1 += __shfl_down(d1, o, 64);
      d2 += __shfl_down(d2, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      red[2 * (threadIdx.x >> 6)] = d1;
      red[2 * (threadIdx.x >> 6) + 1] = d2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double a = 0.0, b = 0.0;
      for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
        a += red[2 * k];
        b += red[2 * k + 1];
      }
      part[0] = a;
      part[1] = b;
    }
  }
}

// one tile of vertices: the residual rows and/or the BCSR block rows it owns.
// MODE fixes the work at compile time (1 residual, 2 Jacobian, 4 apply, 8
// point coefficients), so each entry point gets registers and LDS for its own
// path only; MODE 0 takes it from mode_rt (the batched entry point).  GI,
// GB, GX, GR, GU: the pointer types of the tables, the vectors read and the
// outputs (plain pointers, or global-memory ones where they come from a
// batch member's descriptor).
template <int MODE, typename GI, typename GB, typename GX, typename GR, typename GU>
__device__ __forceinline__ void tile_body(int tile, GI __restrict__ g2l, GB __restrict__ cmask, GX __restrict__ x,
         GX __restrict__ base, double t, double cm, double ca, GR __restrict__ r, GR __restrict__ val,
         GI __restrict__ own_start, GU __restrict__ own_mask, int n0, int n1, int n2, double ox, double oy,
         double oz, double hx, double hy, double hz, int mode_rt, const double* __restrict__ prm) {
  const int mode = MODE ? MODE : mode_rt;
  __shared__ double xs[NVB * BC];
  __shared__ int sl[NVB];
  __shared__ unsigned char cms[NCB];
  __shared__ double work[MODE == 1 ? RL_WORDS : (MODE == 2 || MODE == 4 || MODE == 8) ? JW_WORDS : UNION_WORDS];
  // mode 4 with small blocks: the staged coefficients are applied cell by
  // cell (cells_apply), whose corner sums gather on the vertex box here
  constexpr bool CELL_APPLY = MODE == 4 && BC <= 4 && NCB <= 256;
  __shared__ double acc4[CELL_APPLY ? NVB * BC : 1];
  const int ntx = (n0 + 1 + TX - 1) / TX;
  const int nty = DIM > 1 ? (n1 + 1 + TY - 1) / TY : 1;
  const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
  const int vi0 = tx * TX - 1, vj0 = DIM > 1 ? ty * TY - 1 : 0, vk0 = DIM > 2 ? tz * TZ - 1 : 0;
  // phase 0: vertex box (state, slots) and cell box (class masks)
  for (int b = threadIdx.x; b < NVB; b += blockDim.x) {
    const int bi = b % XB, bj = (b / XB) % YB, bk = b / (XB * YB);
    const int gi = vi0 + bi, gj = vj0 + bj, gk = vk0 + bk;
    int s = -1;
    if (gi >= 0 && gi <= n0 && gj >= 0 && gj <= (DIM > 1 ? n1 : 0) && gk >= 0 && gk <= (DIM > 2 ? n2 : 0))
      s = g2l[gi + (n0 + 1) * (gj + (n1 + 1) * gk)];
    sl[b] = s;
    for (int i = 0; i < BC; ++i) xs[b * BC + i] = s >= 0 ? x[(size_t)s * DCG_B + i] : 0.0;
  }
  for (int c = threadIdx.x; c < NCB; c += blockDim.x) {
    const int ci = c % CXB, cj = (c / CXB) % CYB, ck = c / (CXB * CYB);
    const int gi = vi0 + ci, gj = vj0 + cj, gk = vk0 + ck;
    unsigned char m = 0;
    if (gi >= 0 && gi < n0 && gj >= 0 && gj < (DIM > 1 ? n1 : 1) && gk >= 0 && gk < (DIM > 2 ? n2 : 1))
      m = cmask[gi + n0 * (gj + n1 * gk)];
    cms[c] = m;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  if (mode & 14) {
    // mode 2 writes the BCSR block rows; mode 4 applies them instead (the
    // matrix-free Jacobian: r = J(x) v with the direction v passed as base);
    // mode 8 stores the point coefficients themselves (val) and the rows'
    // diagonal blocks (r)
    double dv[NDV > 0 ? NDV : 1];
    uniform_diffusion(t, prm, dv);
    // phase 1: Jacobian coefficients per (cell, class, point); class and point
    // are uniform per wave so the reference-element tables are scalar loads.
    // Only the structural nonzeros (MCI[n], MCJ[n]) of the point coefficient
    // block are kept; LDS layout is cell-minor: work[((p*NQ + q)*NMC + n)*NCB + cell].
    for (int pq0 = wave; pq0 < NCLASS * NQ; pq0 += nwave) {
      const int pq = __builtin_amdgcn_readfirstlane(pq0);
      const int p = pq / NQ, q = pq % NQ;
      for (int cl = lane; cl < NCB; cl += 64) {
        if (!((cms[cl] >> p) & 1u)) continue;
        const int ci = cl % CXB, cj = (cl / CXB) % CYB, ck = cl / (CXB * CYB);
        // value and gradient at the point, straight from the staged corners
        double u[BC], g[BC][3];
        for (int i = 0; i < BC; ++i) {
          u[i] = 0.0;
          g[i][0] = g[i][1] = g[i][2] = 0.0;
        }
        for (int a = 0; a < NV; ++a) {
          const int c = CORNER[p][a];
          const int b = (ci + (c & 1)) + XB * ((cj + ((c >> 1) & 1)) + YB * (ck + ((c >> 2) & 1)));
          const double pa = PHI[q][a], ga0 = GP(p, q, a, 0);
          const double ga1 = DIM > 1 ? GP(p, q, a, DIM > 1 ? 1 : 0) : 0.0, ga2 = DIM > 2 ? GP(p, q, a, DIM > 2 ? 2 : 0) : 0.0;
          for (int i = 0; i < BC; ++i) {
            const double xv = xs[b * BC + i];
            u[i] += pa * xv;
            g[i][0] += ga0 * xv;
            if (DIM > 1) g[i][1] += ga1 * xv;
            if (DIM > 2) g[i][2] += ga2 * xv;
          }
        }
        const double px = ox + (vi0 + ci) * hx + XQP(p, q, 0);
        const double py = oy + (vj0 + cj) * hy + XQP(p, q, 1);
        const double pz = oz + (vk0 + ck) * hz + XQP(p, q, 2);
        double rs[BC], rf[BC][DIM], mc[NMC > 0 ? NMC : 1];
        qp(q, u, g, px, py, pz, t, cm, ca, rs, rf, mc, nullptr, nullptr, prm, nullptr);
#pragma unroll
        for (int n = 0; n < NMC; ++n) work[(size_t)(pq * NMC + n) * NCB + cl] = mc[n];
      }
    }
    __syncthreads();
    if constexpr (MODE == 8) {
      // the coefficients of the tile's owned cells (lowest corner among its
      // vertices) into the grid-wide array, word-major:
      // val[((p*NQ + q)*NMC + n) * cells of the grid + cell]; cells outside
      // the compartment carry values that are never read
      if constexpr (NMC > 0) {
        const size_t ncg = (size_t)n0 * (size_t)n1 * (size_t)n2;
        for (int k = threadIdx.x; k < OWN_WORDS; k += blockDim.x) {
          const int w = k / NOWN, o = k % NOWN;
          const int oi = o % TX, oj = DIM > 1 ? (o / TX) % TY : 0, ok = DIM > 2 ? o / (TX * TY) : 0;
          const int gi = vi0 + 1 + oi, gj = vj0 + OY + oj, gk = vk0 + OZ + ok;
          if (gi >= n0 || (DIM > 1 && gj >= n1) || (DIM > 2 && gk >= n2)) continue;
          val[(size_t)w * ncg + (size_t)gi + (size_t)n0 * ((size_t)gj + (size_t)n1 * gk)] =
              work[(size_t)w * NCB + (oi + 1) + CXB * ((oj + OY) + CYB * (ok + OZ))];
        }
      }
      // diagonal block of every row (the Jacobi preconditioner's input):
      // r[slot * B * B + i * B + j], the identity on padding species
      constexpr int AI = BC > 4 ? 1 : BC, NR = BC / AI;
      for (int it = threadIdx.x; it < NROW * NR; it += blockDim.x) {
        const int lr = it / NR, i0 = (it % NR) * AI;
        const int li = lr % TX, lj = DIM > 1 ? (lr / TX) % TY : 0, lk = DIM > 2 ? lr / (TX * TY) : 0;
        const int s = sl[(li + 1) + XB * ((lj + OY) + YB * (lk + OZ))];
        if (s < 0) continue;
        double acc[AI][BC];
        block_rows<AI>(0, 0, 0, li, lj, lk, i0, cms, work, dv, ca, acc);
        const auto blk = r + (size_t)s * (DCG_B * DCG_B);
#pragma unroll
        for (int i = 0; i < AI; ++i)
#pragma unroll
          for (int j = 0; j < DCG_B; ++j) blk[(i0 + i) * DCG_B + j] = j < BC ? acc[i][j < BC ? j : 0] : 0.0;
        if (i0 + AI == BC)
          for (int i = BC; i < DCG_B; ++i)
            for (int j = 0; j < DCG_B; ++j) blk[i * DCG_B + j] = i == j ? 1.0 : 0.0;
      }
      return;
    }
    if (mode & 4) {
      // the linearization point is no longer needed: stage the direction
      for (int b = threadIdx.x; b < NVB; b += blockDim.x) {
        const int s = sl[b];
        for (int i = 0; i < BC; ++i) xs[b * BC + i] = s >= 0 ? base[(size_t)s * DCG_B + i] : 0.0;
        if constexpr (CELL_APPLY)
          for (int i = 0; i < BC; ++i) acc4[b * BC + i] = 0.0;
      }
      __syncthreads();
      if constexpr (CELL_APPLY) {
        cells_apply<ApplyTile<TX, TY, TZ>, double>(
            [&](int cl, int wd) { return ((const double*)work)[(size_t)wd * NCB + cl]; }, xs, sl, cms, nullptr, acc4,
            base, r, ca, dv, (GX) nullptr, (GX) nullptr, nullptr);
        return;
      }
    }
    if constexpr (BC <= 4) {
    // phase 2 (blocks up to 4 x 4): SPL lanes per row, split over the last
    // axis of the cells around the row vertex.  A lane reads each of its
    // cells' staged point coefficients once and adds them into all of the
    // row's stencil blocks it touches; every index into the accumulator is a
    // compile-time constant.  Slots along the split axis are the corner bit
    // (2 slots) instead of the offset (3 slots); the two lanes of a row then
    // exchange the shared middle plane through LDS.  The halves are uniform
    // per wave (lanes t and t + 64 hold the same row), so no wave runs both
    // halves' cells.  Each block is written exactly once.
    constexpr int SPL = (DIM > 1 && NROW * 2 <= 256) ? 2 : 1;
    constexpr int NCELL = DIM == 1 ? 2 : DIM == 2 ? 4 : 8;
    constexpr int SA = DIM - 1;  // split axis
    constexpr int SX = 3, SY = DIM > 1 ? ((SPL == 2 && SA == 1) ? 2 : 3) : 1, SZ = DIM > 2 ? (SPL == 2 ? 2 : 3) : 1;
    constexpr int NZ1 = NZB > 0 ? NZB : 1;
    // the row's slots from this lane's cells (half: below / above the row
    // vertex on the split axis)
    auto accumulate = [&](int li, int lj, int lk, int half, double (&acc)[SZ][SY][SX][NZ1]) {
#pragma unroll
      for (int a2 = 0; a2 < SZ; ++a2)
#pragma unroll
        for (int a1 = 0; a1 < SY; ++a1)
#pragma unroll
          for (int a0 = 0; a0 < SX; ++a0)
#pragma unroll
            for (int z = 0; z < NZ1; ++z) acc[a2][a1][a0][z] = 0.0;
#pragma unroll
      for (int e = 0; e < NCELL; ++e) {
        const int ex = e & 1, ey = DIM > 1 ? (e >> 1) & 1 : 0, ez = DIM > 2 ? (e >> 2) & 1 : 0;
        if (SPL == 2 && (SA == 2 ? ez : ey) != half) continue;
        const int cl = (li + ex) + CXB * ((lj + ey) + CYB * (lk + ez));
        const unsigned m = cms[cl];
        if (!m) continue;
        // corner of the row vertex in this cell
        const int cv = (1 - ex) | (DIM > 1 ? (1 - ey) << 1 : 0) | (DIM > 2 ? (1 - ez) << 2 : 0);
#pragma unroll
        for (int p = 0; p < NCLASS; ++p) {
          const int la = LOC[p][cv];
          if (la < 0) continue;
          if (!((m >> p) & 1u)) continue;
          // point by point: the NMC coefficients of point q go into every
          // corner's slot before the next point is read (NMC registers, not
          // NQ x NMC); each slot still sums its points in order
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            double mq[NMC > 0 ? NMC : 1];
#pragma unroll
            for (int n = 0; n < NMC; ++n) mq[n] = work[(size_t)((p * NQ + q) * NMC + n) * NCB + cl];
#pragma unroll
            for (int c = 0; c < NV; ++c) {
              const int cw = CORNER[p][c];
              const int bx = cw & 1, by = (cw >> 1) & 1, bz = (cw >> 2) & 1;
              const int sx = bx - (cv & 1) + 1;
              const int sy = DIM > 1 ? ((SPL == 2 && SA == 1) ? by : by - ((cv >> 1) & 1) + 1) : 0;
              const int sz = DIM > 2 ? (SPL == 2 ? bz : bz - ((cv >> 2) & 1) + 1) : 0;
              double* ac = acc[sz][sy][sx];
              const double w = PP[q][la][c];
#pragma unroll
              for (int n = 0; n < NMC; ++n) ac[MCZ[n]] += w * mq[n];
            }
          }
#pragma unroll
          for (int c = 0; c < NV; ++c) {
            const int cw = CORNER[p][c];
            const int bx = cw & 1, by = (cw >> 1) & 1, bz = (cw >> 2) & 1;
            const int sx = bx - (cv & 1) + 1;
            const int sy = DIM > 1 ? ((SPL == 2 && SA == 1) ? by : by - ((cv >> 1) & 1) + 1) : 0;
            const int sz = DIM > 2 ? (SPL == 2 ? bz : bz - ((cv >> 2) & 1) + 1) : 0;
            double* ac = acc[sz][sy][sx];
            if constexpr (CONST_DIFFUSION) {
#pragma unroll
              for (int n = 0; n < NKZ; ++n) ac[KZZ[n]] += ca * kd_rt(p, la, c, KZI[n], KZJ[n]);
            } else {
#pragma unroll
              for (int n = 0; n < NDV; ++n) ac[DVZ[n]] += ca * dv[n] * KDU[p][la][c][DVK[n]][DVL[n]];
            }
          }
        }
      }
    };
    // stencil offset of slot (a0, a1, a2) of a lane; false when the slot is
    // the other lane's (the lower lane writes the middle plane)
    auto slot_offset = [&](int a0, int a1, int a2, int half, int& dx, int& dy, int& dz) -> bool {
      dx = a0 - 1;
      dy = DIM > 1 ? a1 - 1 : 0;
      dz = DIM > 2 ? a2 - 1 : 0;
      if constexpr (SPL == 2 && SA == 2) {
        dz = a2 - (1 - half);
        return half == 0 || a2 == 1;
      }
      if constexpr (SPL == 2 && SA == 1) {
        dy = a1 - (1 - half);
        return half == 0 || a1 == 1;
      }
      return true;
    };
    // r = J v over the row (v staged in xs), summed over this lane's slots
    auto apply_row = [&](int vb, int half, const double (&acc)[SZ][SY][SX][NZ1], double (&ya)[BC]) {
#pragma unroll
      for (int i = 0; i < BC; ++i) ya[i] = 0.0;
#pragma unroll
      for (int a2 = 0; a2 < SZ; ++a2)
#pragma unroll
        for (int a1 = 0; a1 < SY; ++a1)
#pragma unroll
          for (int a0 = 0; a0 < SX; ++a0) {
            int dx, dy, dz;
            slot_offset(a0, a1, a2, half, dx, dy, dz);
            const int nb = vb + dx + XB * (dy + YB * dz);
#pragma unroll
            for (int i = 0; i < BC; ++i)
#pragma unroll
              for (int j = 0; j < BC; ++j) {
                const int z = ZIX[i][j];
                if (z >= 0) ya[i] += acc[a2][a1][a0][z >= 0 ? z : 0] * xs[nb * BC + j];
              }
          }
    };
    auto write_row = [&](int s, int half, const double (&acc)[SZ][SY][SX][NZ1]) {
      const unsigned mk = own_mask[s];
      const int st = own_start[s];
#pragma unroll
      for (int a2 = 0; a2 < SZ; ++a2)
#pragma unroll
        for (int a1 = 0; a1 < SY; ++a1)
#pragma unroll
          for (int a0 = 0; a0 < SX; ++a0) {
            int dx, dy, dz;
            if (!slot_offset(a0, a1, a2, half, dx, dy, dz)) continue;
            const int o = (dx + 1) + (DIM > 1 ? 3 * (dy + 1) : 0) + (DIM > 2 ? 9 * (dz + 1) : 0);
            if (!((mk >> o) & 1u)) continue;
            const auto blk = val + (size_t)(st + __popc(mk & ((1u << o) - 1u))) * (DCG_B * DCG_B);
            const bool diag = o == NST / 2;
#pragma unroll
            for (int i = 0; i < DCG_B; ++i)
#pragma unroll
              for (int j = 0; j < DCG_B; ++j) {
                const int z = (i < BC && j < BC) ? ZIX[i < BC ? i : 0][j < BC ? j : 0] : -1;
                blk[i * DCG_B + j] = z >= 0 ? acc[a2][a1][a0][z >= 0 ? z : 0] : ((diag && i == j && i >= BC) ? 1.0 : 0.0);
              }
          }
    };
    if constexpr (SPL == 2) {
      const int tid = threadIdx.x;
      const int half = (tid >> 6) & 1;
      const int lr = (tid >> 7) * 64 + (tid & 63);
      const int li = lr % TX, lj = (lr / TX) % TY, lk = DIM > 2 ? lr / (TX * TY) : 0;
      const int vb = (li + 1) + XB * ((lj + OY) + YB * (lk + OZ));
      const int s = lr < NROW ? sl[vb] : -1;
      double acc[SZ][SY][SX][NZ1];
      if (s >= 0) accumulate(li, lj, lk, half, acc);
      if (mode & 4) {
        // the apply is linear: each lane applies its own slots (its part of
        // the middle plane included) and the two sums are added
        double ya[BC];
        if (s >= 0) apply_row(vb, half, acc, ya);
        __syncthreads();  // every lane is done with the staged coefficients
        if (s >= 0 && half == 1)
          for (int i = 0; i < BC; ++i) work[lr * BC + i] = ya[i];
        __syncthreads();
        if (s >= 0 && half == 0)
          for (int i = 0; i < DCG_B; ++i)
            r[(size_t)s * DCG_B + i] = i < BC ? ya[i < BC ? i : 0] + work[lr * BC + (i < BC ? i : 0)]
                                              : base[(size_t)s * DCG_B + i];
      } else {
        // the middle plane (corner bit 1 below, 0 above) in rounds of XR
        // slots that fit the coefficient array
        constexpr int MSL = SA == 2 ? SY * SX : SX;
        constexpr int XR0 = JW_WORDS / (NROW * NZ1);
        constexpr int XR = XR0 < 1 ? 1 : XR0 < MSL ? XR0 : MSL;
#pragma unroll
        for (int k0 = 0; k0 < MSL; k0 += XR) {
          __syncthreads();  // the coefficients (or the last round) are read
          if (s >= 0 && half == 1) {
#pragma unroll
            for (int k = k0; k < (k0 + XR < MSL ? k0 + XR : MSL); ++k)
#pragma unroll
              for (int z = 0; z < NZ1; ++z)
                work[((k - k0) * NROW + lr) * NZ1 + z] = SA == 2 ? acc[0][k / SX][k % SX][z] : acc[0][0][k][z];
          }
          __syncthreads();
          if (s >= 0 && half == 0) {
#pragma unroll
            for (int k = k0; k < (k0 + XR < MSL ? k0 + XR : MSL); ++k)
#pragma unroll
              for (int z = 0; z < NZ1; ++z) {
                const double v = work[((k - k0) * NROW + lr) * NZ1 + z];
                if constexpr (SA == 2) acc[1][k / SX][k % SX][z] += v;
                else acc[0][1][k][z] += v;
              }
          }
        }
        if (s >= 0) write_row(s, half, acc);
      }
    } else {
      for (int lr = threadIdx.x; lr < NROW; lr += blockDim.x) {
        const int li = lr % TX, lj = DIM > 1 ? (lr / TX) % TY : 0, lk = DIM > 2 ? lr / (TX * TY) : 0;
        const int vb = (li + 1) + XB * ((lj + OY) + YB * (lk + OZ));
        const int s = sl[vb];
        if (s < 0) continue;
        double acc[SZ][SY][SX][NZ1];
        accumulate(li, lj, lk, 0, acc);
        if (mode & 4) {
          double ya[BC];
          apply_row(vb, 0, acc, ya);
          for (int i = 0; i < DCG_B; ++i) r[(size_t)s * DCG_B + i] = i < BC ? ya[i < BC ? i : 0] : base[(size_t)s * DCG_B + i];
        } else {
          write_row(s, 0, acc);
        }
      }
    }
    } else {
    // phase 2 (blocks above 4 x 4): one BCSR block per (row, stencil offset),
    // written once; the offset is uniform per wave and a lane builds one block
    // row, so the accumulator stays at BC registers and small tiles still
    // fill a wave.
    constexpr int RL = DCG_B;  // lanes per row
    if (mode & 4) {
      // r = J v: a lane keeps its block rows' sums over all stencil offsets
      // (the offset loop is uniform across the workgroup)
      constexpr int PER = (NROW * RL + 255) / 256;
      double ya[PER];
#pragma unroll
      for (int k2 = 0; k2 < PER; ++k2) ya[k2] = 0.0;
      for (int o = 0; o < NST; ++o) {
        const int dx = o % 3 - 1, dy = DIM > 1 ? (o / 3) % 3 - 1 : 0, dz = DIM > 2 ? o / 9 - 1 : 0;
#pragma unroll
        for (int k2 = 0; k2 < PER; ++k2) {
          const int it = threadIdx.x + 256 * k2;
          if (it >= NROW * RL) continue;
          const int lr = it / RL, ib = it % RL;
          const int li = lr % TX, lj = DIM > 1 ? (lr / TX) % TY : 0, lk = DIM > 2 ? lr / (TX * TY) : 0;
          const int vb = (li + 1) + XB * ((lj + OY) + YB * (lk + OZ));
          if (sl[vb] < 0 || ib >= BC) continue;
          double acc[1][BC];
          block_rows<1>(dx, dy, dz, li, lj, lk, ib, cms, work, dv, ca, acc);
          const int nb = vb + dx + XB * (dy + YB * dz);
#pragma unroll
          for (int j = 0; j < BC; ++j) ya[k2] += acc[0][j] * xs[nb * BC + j];
        }
      }
#pragma unroll
      for (int k2 = 0; k2 < PER; ++k2) {
        const int it = threadIdx.x + 256 * k2;
        if (it >= NROW * RL) continue;
        const int lr = it / RL, ib = it % RL;
        const int li = lr % TX, lj = DIM > 1 ? (lr / TX) % TY : 0, lk = DIM > 2 ? lr / (TX * TY) : 0;
        const int s = sl[(li + 1) + XB * ((lj + OY) + YB * (lk + OZ))];
        if (s < 0) continue;
        r[(size_t)s * DCG_B + ib] = ib < BC ? ya[k2] : base[(size_t)s * DCG_B + ib];
      }
    } else
    for (int o0 = wave; o0 < NST; o0 += nwave) {
      const int o = __builtin_amdgcn_readfirstlane(o0);
      const int dx = o % 3 - 1, dy = DIM > 1 ? (o / 3) % 3 - 1 : 0, dz = DIM > 2 ? o / 9 - 1 : 0;
      const bool diag = o == NST / 2;
      for (int it = lane; it < NROW * RL; it += 64) {
        const int lr = it / RL, ib = it % RL;
        const int li = lr % TX, lj = DIM > 1 ? (lr / TX) % TY : 0, lk = DIM > 2 ? lr / (TX * TY) : 0;
        const int vb = (li + 1) + XB * ((lj + OY) + YB * (lk + OZ));
        const int s = sl[vb];
        if (s < 0) continue;
        const unsigned mk = own_mask[s];
        if (!((mk >> o) & 1u)) continue;
        const int pos = own_start[s] + __popc(mk & ((1u << o) - 1u));
        constexpr int AI = BC > 4 ? 1 : BC;  // block rows per lane
        double acc[AI][BC];
        const int i0 = BC > 4 ? ib : 0;  // first block row of this lane
        block_rows<AI>(dx, dy, dz, li, lj, lk, i0, cms, work, dv, ca, acc);
        const auto blk = val + (size_t)pos * (DCG_B * DCG_B);
#pragma unroll
        for (int i = 0; i < (BC > 4 ? 1 : DCG_B); ++i) {
          const int bi = i0 + i;
#pragma unroll
          for (int j = 0; j < DCG_B; ++j)
            blk[bi * DCG_B + j] = (bi < BC && j < BC) ? acc[i < AI ? i : 0][j < BC ? j : 0] : ((diag && bi == j) ? 1.0 : 0.0);
        }
      }
    }
    }
    __syncthreads();
  }
  // Box cells with blocks of up to 4 species keep a cell's corner sums in
  // registers and store them once, component-major (corner c, species i of
  // box cell cl at work[(c * BC + i) * NCB + cl]), so the lanes of a wave
  // (consecutive cells) store and the rows read consecutive words; larger
  // blocks and simplex classes add into a per-cell array whose odd stride
  // (RL_CELL doubles) spreads the lanes over the LDS banks
  // (work[cl * RL_CELL + c * BC + i]).
  constexpr bool RES_REG = BC <= 4 && NCLASS == 1;
  if (mode & 1) {
    // phase 1: element residual per cell, folded onto the cell's corners
    for (int cl = threadIdx.x; cl < NCB; cl += blockDim.x) {
      const unsigned m = cms[cl];
      const int ci = cl % CXB, cj = (cl / CXB) % CYB, ck = cl / (CXB * CYB);
      const double cx = ox + (vi0 + ci) * hx, cy = oy + (vj0 + cj) * hy, cz = oz + (vk0 + ck) * hz;
      if constexpr (RES_REG) {
        double oc[NCORNER][BC];
#pragma unroll
        for (int c = 0; c < NCORNER; ++c)
#pragma unroll
          for (int i = 0; i < BC; ++i) oc[c][i] = 0.0;
        if (m) {
#pragma unroll
          for (int p = 0; p < NCLASS; ++p) {
            if (!((m >> p) & 1u)) continue;
            double xl[NV][BC], rl[NV][BC];
#pragma unroll
            for (int a = 0; a < NV; ++a) {
              const int c = CORNER[p][a];
              const int b = (ci + (c & 1)) + XB * ((cj + ((c >> 1) & 1)) + YB * (ck + ((c >> 2) & 1)));
#pragma unroll
              for (int i = 0; i < BC; ++i) xl[a][i] = xs[b * BC + i];
            }
            residual_rt(p, xl, cx, cy, cz, t, cm, ca, rl, prm);
#pragma unroll
            for (int a = 0; a < NV; ++a)
#pragma unroll
              for (int i = 0; i < BC; ++i) oc[CORNER[p][a]][i] += rl[a][i];
          }
        }
#pragma unroll
        for (int c = 0; c < NCORNER; ++c)
#pragma unroll
          for (int i = 0; i < BC; ++i) work[(size_t)(c * BC + i) * NCB + cl] = oc[c][i];
      } else {
        double* dst = work + (size_t)cl * RL_CELL;
        for (int k = 0; k < NCORNER * BC; ++k) dst[k] = 0.0;
        if (!m) continue;
        for (int p = 0; p < NCLASS; ++p) {
          if (!((m >> p) & 1u)) continue;
          if constexpr (BC > 4) {
            int bidx[NV];
            for (int a = 0; a < NV; ++a) {
              const int c = CORNER[p][a];
              bidx[a] = (ci + (c & 1)) + XB * ((cj + ((c >> 1) & 1)) + YB * (ck + ((c >> 2) & 1)));
            }
            residual_lds(p, xs, bidx, cx, cy, cz, t, cm, ca, dst, prm);
          } else {
            double xl[NV][BC], rl[NV][BC];
            for (int a = 0; a < NV; ++a) {
              const int c = CORNER[p][a];
              const int b = (ci + (c & 1)) + XB * ((cj + ((c >> 1) & 1)) + YB * (ck + ((c >> 2) & 1)));
              for (int i = 0; i < BC; ++i) xl[a][i] = xs[b * BC + i];
            }
            residual_rt(p, xl, cx, cy, cz, t, cm, ca, rl, prm);
            for (int a = 0; a < NV; ++a)
              for (int i = 0; i < BC; ++i) dst[CORNER[p][a] * BC + i] += rl[a][i];
          }
        }
      }
    }
    __syncthreads();
    // phase 2: r = base + sum of the cells around each row vertex
    for (int it = threadIdx.x; it < NROW * DCG_B; it += blockDim.x) {
      const int lr = it / DCG_B, i = it % DCG_B;
      const int li = lr % TX, lj = DIM > 1 ? (lr / TX) % TY : 0, lk = DIM > 2 ? lr / (TX * TY) : 0;
      const int vb = (li + 1) + XB * ((lj + OY) + YB * (lk + OZ));
      const int s = sl[vb];
      if (s < 0) continue;
      double v = base ? base[(size_t)s * DCG_B + i] : 0.0;
      if (i < BC) {
        for (int ez = 0; ez < (DIM > 2 ? 2 : 1); ++ez)
          for (int ey = 0; ey < (DIM > 1 ? 2 : 1); ++ey)
            for (int ex = 0; ex < 2; ++ex) {
              const int cl = (li + ex) + CXB * ((lj + ey) + CYB * (lk + ez));
              if (!cms[cl]) continue;
              const int cv = (1 - ex) | (DIM > 1 ? (1 - ey) << 1 : 0) | (DIM > 2 ? (1 - ez) << 2 : 0);
              v += RES_REG ? work[(size_t)(cv * BC + i) * NCB + cl] : work[(size_t)cl * RL_CELL + cv * BC + i];
            }
      }
      r[(size_t)s * DCG_B + i] = v;
    }
  }
}

// y = J v from stored point coefficients (mode 8 of tile_body: each cell's
// coefficients once, in the grid-wide word-major array pc[word * cells +
// cell]): per cell of the tile's box, v at its corners is interpolated to
// each point, multiplied by the point's coefficient block and tested against
// the basis functions, the uniform diffusion adds the element stiffness
// times v; the cells' corner sums are folded onto the tile's rows as in the
// residual.  T is the value type of pc, v and y (double, or float inside the
// refinement loop).  tca: the time and stiffness weight of the
// linearisation, written to device memory by the set-up, so an apply's
// arguments are the same for every linearisation.  g2l null: the
// compartment's slots are the grid's vertices in order.  With part, the
// workgroup also writes its rows' partial dots (y.w1, y.w2) to part[0] and
// part[1] (w2 may be null), as the BiCGSTAB finaliser reads.
// (pointers typed in the global address space: the batched entry takes them
// from member descriptors, where the compiler could not tell their space
// and would use flat accesses)
template <typename X>
using dcg_gp = __attribute__((address_space(1))) X*;
template <typename TL, typename T>
__device__ __forceinline__ void pc_apply(int tile, dcg_gp<const int> __restrict__ g2l,
                                         dcg_gp<const unsigned char> __restrict__ cmask,
                                         dcg_gp<const T> __restrict__ pc, dcg_gp<const T> __restrict__ v,
                                         dcg_gp<T> __restrict__ y, dcg_gp<const double> __restrict__ tca, int n0,
                                         int n1, int n2, const double* __restrict__ prm,
                                         dcg_gp<const T> __restrict__ w1, dcg_gp<const T> __restrict__ w2,
                                         double* __restrict__ part) {
  constexpr int TX = TL::TX, TY = TL::TY, TZ = TL::TZ, XB = TL::XB, YB = TL::YB, CXB = TL::CXB, CYB = TL::CYB;
  constexpr int NVB = TL::NVB, NCB = TL::NCB;
  const double t = tca[0], ca = tca[1];
  // corner values stay in registers for small blocks, in LDS otherwise; with
  // small blocks and one cell per thread the corner sums go straight onto
  // the vertices (SCATTER), otherwise through a per-cell corner array
  constexpr bool VREG = BC <= 4;
  constexpr bool SCATTER = VREG && NCB <= TL::NT;
  __shared__ T vs[NVB * BC];
  __shared__ int sl[NVB];
  __shared__ unsigned char cms[NCB];
  __shared__ size_t coff[NCB];
  __shared__ T out[SCATTER ? 1 : NCB * NCORNER * BC];
  __shared__ T accv[SCATTER ? NVB * BC : 1];
  const int ntx = (n0 + 1 + TX - 1) / TX;
  const int nty = DIM > 1 ? (n1 + 1 + TY - 1) / TY : 1;
  const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
  const int vi0 = tx * TX - 1, vj0 = DIM > 1 ? ty * TY - 1 : 0, vk0 = DIM > 2 ? tz * TZ - 1 : 0;
  const size_t WS = (size_t)n0 * (size_t)n1 * (size_t)n2;
  for (int b = threadIdx.x; b < NVB; b += blockDim.x) {
    const int bi = b % XB, bj = (b / XB) % YB, bk = b / (XB * YB);
    const int gi = vi0 + bi, gj = vj0 + bj, gk = vk0 + bk;
    int s = -1;
    if (gi >= 0 && gi <= n0 && gj >= 0 && gj <= (DIM > 1 ? n1 : 0) && gk >= 0 && gk <= (DIM > 2 ? n2 : 0)) {
      const int vid = gi + (n0 + 1) * (gj + (n1 + 1) * gk);
      s = g2l ? g2l[vid] : vid;
    }
    sl[b] = s;
    for (int i = 0; i < BC; ++i) vs[b * BC + i] = s >= 0 ? v[(size_t)s * DCG_B + i] : T(0);
    if constexpr (SCATTER)
      for (int i = 0; i < BC; ++i) accv[b * BC + i] = T(0);
  }
  for (int c = threadIdx.x; c < NCB; c += blockDim.x) {
    const int ci = c % CXB, cj = (c / CXB) % CYB, ck = c / (CXB * CYB);
    const int gi = vi0 + ci, gj = vj0 + cj, gk = vk0 + ck;
    unsigned char m = 0;
    size_t off = 0;
    if (gi >= 0 && gi < n0 && gj >= 0 && gj < (DIM > 1 ? n1 : 1) && gk >= 0 && gk < (DIM > 2 ? n2 : 1)) {
      off = (size_t)gi + (size_t)n0 * ((size_t)gj + (size_t)n1 * gk);
      m = cmask[off];
    }
    cms[c] = m;
    coff[c] = off;
  }
  __syncthreads();
  double dv[NDV > 0 ? NDV : 1];
  uniform_diffusion(t, prm, dv);
  cells_apply<TL, T>(
      [&](int cl, int wd) -> T { return pc[coff[cl] + (size_t)wd * WS]; },
      vs, sl, cms, out, accv, v, y, ca, dv, w1, w2, part);
}

}  // namespace %C%

// one entry point per kind of work, same arguments: _r residual (r = base +
// local residuals), _j Jacobian block rows into val, _a r = J(x) base
#define DCG_TILE_ENTRY(NAME, M)                                                                                   \
  extern "C" __global__ void __launch_bounds__(256)                                                               \
  NAME(const int* __restrict__ tiles, const int* __restrict__ g2l, const unsigned char* __restrict__ cmask,         \
       const double* __restrict__ x, const double* __restrict__ base, double t, double cm, double ca,              \
       double* __restrict__ r, double* __restrict__ val, const int* __restrict__ own_start,                        \
       const unsigned* __restrict__ own_mask, int n0, int n1, int n2, double ox, double oy, double oz, double hx,  \
       double hy, double hz, int mode, const double* __restrict__ prm, const int* __restrict__ stop) {             \
    /* stop: a Krylov solve's done flag (the matrix-free apply inside BiCGSTAB) */                                \
    if (stop && *stop) return;                                                                                    \
    %C%::tile_body<M>(tiles[blockIdx.x], g2l, cmask, x, base, t, cm, ca, r, val, own_start, own_mask, n0, n1, n2,  \
                      ox, oy, oz, hx, hy, hz, mode, prm);                                                         \
  }
DCG_TILE_ENTRY(%KERNEL%_r, 1)
DCG_TILE_ENTRY(%KERNEL%_j, 2)
DCG_TILE_ENTRY(%KERNEL%_a, 4)
#undef DCG_TILE_ENTRY
)";

// Entry points of the point-coefficient Jacobian (emitted on request): _q
// stores the point coefficients at x into val (per cell and point) and the
// rows' diagonal blocks into r; _p / _pf apply them, y = J v, in FP64 / FP32.
const char* kTiledPoints = R"(
extern "C" __global__ void __launch_bounds__(256)
%KERNEL%_q(const int* __restrict__ tiles, const int* __restrict__ g2l, const unsigned char* __restrict__ cmask,
           const double* __restrict__ x, const double* __restrict__ base, double t, double cm, double ca,
           double* __restrict__ r, double* __restrict__ val, const int* __restrict__ own_start,
           const unsigned* __restrict__ own_mask, int n0, int n1, int n2, double ox, double oy, double oz, double hx,
           double hy, double hz, int mode, const double* __restrict__ prm, const int* __restrict__ stop) {
  (void)stop;
  %C%::tile_body<8>(tiles[blockIdx.x], g2l, cmask, x, base, t, cm, ca, r, val, own_start, own_mask, n0, n1, n2, ox,
                    oy, oz, hx, hy, hz, mode, prm);
}
// The grid is a multiple of 8 workgroups; workgroup b takes list position
// (b % 8) * (grid / 8) + b / 8, so each XCD (workgroups are dealt to the 8
// XCDs in turn) walks one contiguous run of tiles and finds the neighbour
// tiles' coefficients in its own L2.  Position c writes part[2c], part[2c+1].
#define DCG_POINT_ENTRY(NAME, T, TL)                                                                              \
  extern "C" __global__ void __launch_bounds__(TL::NT, TL::MINW)                                                  \
  NAME(const int* __restrict__ tiles, int ntile, const int* __restrict__ g2l,                                     \
       const unsigned char* __restrict__ cmask, const T* __restrict__ pc, const T* __restrict__ v,                 \
       T* __restrict__ y, const double* __restrict__ tca, int n0, int n1, int n2, const double* __restrict__ prm,  \
       const int* __restrict__ stop, const T* __restrict__ w1, const T* __restrict__ w2,                          \
       double* __restrict__ part) {                                                                               \
    if (stop && *stop) return;                                                                                    \
    const int c = (int)(blockIdx.x % 8u) * (int)(gridDim.x / 8u) + (int)(blockIdx.x / 8u);                       \
    if (c >= ntile) return;                                                                                       \
    %C%::pc_apply<TL, T>(tiles[c], (%C%::dcg_gp<const int>)g2l, (%C%::dcg_gp<const unsigned char>)cmask,          \
                         (%C%::dcg_gp<const T>)pc, (%C%::dcg_gp<const T>)v, (%C%::dcg_gp<T>)y,                    \
                         (%C%::dcg_gp<const double>)tca, n0, n1, n2, prm, (%C%::dcg_gp<const T>)w1,               \
                         (%C%::dcg_gp<const T>)w2, part ? part + 2 * (size_t)c : nullptr);                        \
  }
%POINT_ENTRIES%
#undef DCG_POINT_ENTRY
)";

// Batched point-coefficient Jacobian (members with point storage): _qbatch
// stores every listed member's coefficients and diagonal blocks at the
// member's x (as _q does) and records its (t, ca) for the applies; _pbatch /
// _pfbatch apply them over (member, tile) work items with the member's
// vectors by selector, writing item c's partial dots to part[2 (off + c)],
// and skip members whose solve (skip 0) or refinement (skip 1) is done.
const char* kTiledPointBatch = R"(
extern "C" __global__ void __launch_bounds__(256)
%KERNEL%_qbatch(const int2* __restrict__ work, const dcg_asm_member* __restrict__ mem,
                const double* __restrict__ coef, const int* __restrict__ act, int xsel) {
  const int2 w = work[blockIdx.x];
  if (act && !act[w.x]) return;
  const double* cf = coef + 4 * (size_t)w.x;
  if (cf[3] == 0.0) return;
  const dcg_asm_member& M = mem[w.x];
  if (threadIdx.x == 0) {
    // every workgroup of the member writes the same two values
    M.tca[0] = cf[0];
    M.tca[1] = cf[2];
  }
  // (the member's pointers in the global address space: global, not flat, accesses)
  using GX = %C%::dcg_gp<const double>;
  using GR = %C%::dcg_gp<double>;
  %C%::tile_body<8>(w.y, (%C%::dcg_gp<const int>)M.g2l[%COMP%], (%C%::dcg_gp<const unsigned char>)M.cmask[%COMP%],
                    (GX)M.vec[xsel], (GX) nullptr, cf[0], cf[1], cf[2], (GR)M.pdiag, (GR)M.pc[%COMP%],
                    (%C%::dcg_gp<const int>)M.own_start, (%C%::dcg_gp<const unsigned>)M.own_mask, M.n[0], M.n[1],
                    M.n[2], M.origin[0], M.origin[1], M.origin[2], M.h[0], M.h[1], M.h[2], 8, M.prm);
}
#define DCG_PBATCH_ENTRY(NAME, T, VEC, PC, TL)                                                                    \
  extern "C" __global__ void __launch_bounds__(TL::NT, TL::MINW)                                                  \
  NAME(const int2* __restrict__ work, int nwork, int part_off, const dcg_pc_member* __restrict__ mem, int xsel,   \
       int ysel, int w1sel, int w2sel, int skip, double* __restrict__ part) {                                      \
    const int c = (int)(blockIdx.x % 8u) * (int)(gridDim.x / 8u) + (int)(blockIdx.x / 8u);                       \
    if (c >= nwork) return;                                                                                       \
    const int2 w = work[c];                                                                                       \
    const dcg_pc_member& M = mem[w.x];                                                                            \
    if (skip == 0 && *M.done) return;                                                                             \
    if (skip == 1 && *M.irdone) return;                                                                           \
    using GT = %C%::dcg_gp<const T>;                                                                              \
    %C%::pc_apply<TL, T>(w.y, (%C%::dcg_gp<const int>)M.g2l, (%C%::dcg_gp<const unsigned char>)M.cmask,          \
                         (GT)M.PC, (GT)M.VEC[xsel], (%C%::dcg_gp<T>)M.VEC[ysel],                                  \
                         (%C%::dcg_gp<const double>)M.tca, M.n[0], M.n[1], M.n[2], M.prm,                         \
                         w1sel >= 0 ? (GT)M.VEC[w1sel] : nullptr, w2sel >= 0 ? (GT)M.VEC[w2sel] : nullptr,        \
                         part ? part + 2 * (size_t)(part_off + c) : nullptr);                                     \
  }
%PBATCH_ENTRIES%
#undef DCG_PBATCH_ENTRY
)";

// Batched form: one launch over (member, tile) pairs of many models that share
// this source; each member's tables, spacing, parameters and vectors come from
// its descriptor, its time and mass / stiffness weights from coef[4 m .. 4 m + 3]
// = (t, cm, ca, on).  Members with act[m] == 0 or on == 0 are skipped.
const char* kTiledBatch = R"(
extern "C" __global__ void __launch_bounds__(256)
%KERNEL%_batch(const int2* __restrict__ work, const dcg_asm_member* __restrict__ mem,
               const double* __restrict__ coef, const int* __restrict__ act, int xsel, int bsel, int rsel, int mode) {
  const int2 w = work[blockIdx.x];
  if (act && !act[w.x]) return;
  const double* cf = coef + 4 * (size_t)w.x;
  if (cf[3] == 0.0) return;
  const dcg_asm_member& M = mem[w.x];
  // (the member's pointers in the global address space: global, not flat, accesses)
  using GX = %C%::dcg_gp<const double>;
  using GR = %C%::dcg_gp<double>;
  %C%::tile_body<0>(w.y, (%C%::dcg_gp<const int>)M.g2l[%COMP%], (%C%::dcg_gp<const unsigned char>)M.cmask[%COMP%],
                    (GX)M.vec[xsel], bsel >= 0 ? (GX)M.vec[bsel] : (GX) nullptr, cf[0], cf[1], cf[2],
                    rsel >= 0 ? (GR)M.vec[rsel] : (GR) nullptr, (mode & 2) ? (GR)M.val : (GR) nullptr,
                    (%C%::dcg_gp<const int>)M.own_start, (%C%::dcg_gp<const unsigned>)M.own_mask, M.n[0], M.n[1],
                    M.n[2], M.origin[0], M.origin[1], M.origin[2], M.h[0], M.h[1], M.h[2], mode, M.prm);
}
)";

}  // namespace

int apply_tile_threads(int dim, std::array<int, 3> t, int species) {
  // one cell of the cell box per thread (corner sums scattered in LDS) for
  // blocks of up to 4 species; larger blocks loop over the cells
  int ncb = t[0] + 1;
  if (dim > 1) ncb *= t[1] + 1;
  if (dim > 2) ncb *= t[2] + 1;
  if (species > 4 || ncb <= 256 || ncb > 1024) return 256;
  return ncb <= 512 ? 512 : 1024;
}

std::vector<std::array<int, 3>> apply_tile_candidates(const Grid& g, int species, std::array<int, 3> setup_tile,
                                                      std::vector<int>* waves) {
  std::vector<std::array<int, 3>> out{setup_tile};
  if (waves) waves->assign(1, 1);
  if (species > 4) return out;  // corner values in LDS: the assembly tile only
  // cell boxes of 256, 512 and 1024 cells around the rows (one cell per
  // thread): squarer boxes re-read fewer halo cells, slabs along x give
  // longer contiguous runs; a few shapes again with registers bounded for
  // more waves per SIMD (more workgroups resident to overlap one's barriers
  // with another's loads)
  std::vector<std::array<int, 3>> cand;
  std::vector<int> wv;
  if (g.dim == 2) {
    cand = {{15, 15, 1}, {31, 7, 1}, {16, 8, 1}, {31, 15, 1}, {31, 31, 1}, {15, 15, 1}, {31, 15, 1}};
    wv = {1, 1, 1, 1, 1, 8, 8};
  } else if (g.dim == 3) {
    cand = {{8, 4, 4}, {7, 7, 3}, {6, 5, 5}, {15, 3, 3}, {7, 7, 7}, {15, 7, 3}, {15, 7, 7}, {7, 7, 7}, {8, 4, 4}};
    wv = {1, 1, 1, 1, 1, 1, 1, 6, 6};  // (8 waves spill in 3D)
  }
  for (size_t i = 0; i < cand.size(); ++i)
    if (cand[i] != setup_tile || wv[i] != 1) {
      out.push_back(cand[i]);
      if (waves) waves->push_back(wv[i]);
    }
  return out;
}

bool tile_eligible(const ModelSpec& m, int comp) {
  if (m.jacobian == JacobianType::Numerical && !m.is_linear) return false;
  for (int s : m.comp_species[comp]) {
    const auto& sp = m.species[s];
    for (int k = 0; k < 3; ++k)
      if (sp.velocity[k]) return false;
    for (const auto& dt : sp.diffusion) {
      if (!dt.jac.empty() || !tensor_uniform(dt.D)) return false;
    }
  }
  return true;
}

std::array<int, 3> choose_tile(const Grid& g, int bc, int nmc) {
  // largest candidate whose LDS (staging + coefficient union) stays under 64 KiB
  std::vector<std::array<int, 3>> cand;
  if (g.dim == 3) cand = {{8, 4, 4}, {4, 4, 4}, {4, 4, 2}, {4, 2, 2}, {2, 2, 2}};
  else if (g.dim == 2) cand = {{16, 8, 1}, {8, 8, 1}, {8, 4, 1}, {4, 4, 1}, {2, 2, 1}};
  else cand = {{128, 1, 1}, {64, 1, 1}, {32, 1, 1}};
  const int nq = static_cast<int>(element_rule(g.dim, g.cube).w.size());
  for (auto c : cand) {
    long xb = c[0] + 2, yb = g.dim > 1 ? c[1] + 2 : 1, zb = g.dim > 2 ? c[2] + 2 : 1;
    long cxb = c[0] + 1, cyb = g.dim > 1 ? c[1] + 1 : 1, czb = g.dim > 2 ? c[2] + 1 : 1;
    long ncb = cxb * cyb * czb;
    long mc = ncb * g.nclass * nq * std::max(1, nmc), rl = ncb * (g.ncorner * bc + 1);
    long bytes = 8 * std::max(mc, rl) + 8 * xb * yb * zb * bc + 4 * xb * yb * zb + ncb;
    if (bytes <= 64 * 1024) return c;
  }
  return cand.back();
}

std::string gen_tiled(const ModelSpec& m, const Grid& g, int comp, const std::string& ns, const std::string& kname,
                      std::array<int, 3> tile, bool geo, bool points,
                      const std::vector<std::array<int, 3>>& apply_tiles, const std::vector<int>& apply_waves) {
  const auto T = element_tables(geo ? unit_cell(g) : g);
  std::ostringstream o;
  o << "namespace " << ns << " {\n";
  o << "constexpr int TX = " << tile[0] << ", TY = " << tile[1] << ", TZ = " << tile[2] << ";\n";
  std::vector<double> pp;
  for (int q = 0; q < T.nq; ++q)
    for (int a = 0; a < T.nv; ++a)
      for (int c = 0; c < T.nv; ++c) pp.push_back(T.phi[q * T.nv + a] * T.phi[q * T.nv + c]);
  o << "__device__ constexpr double PP[NQ][NV][NV] = " << array_init(pp) << ";\n";
  std::vector<int> loc;
  for (int p = 0; p < g.nclass; ++p)
    for (int c = 0; c < g.ncorner; ++c) {
      int l = -1;
      for (int a = 0; a < g.nvpe; ++a)
        if (g.corner[p][a] == c) l = a;
      loc.push_back(l);
    }
  o << "__device__ constexpr int LOC[NCLASS][NCORNER] = " << int_init(loc) << ";\n";
  // structural nonzeros of the point coefficient block mc (mass and reaction
  // Jacobian entries the point function writes)
  const auto nz = mc_pattern(m, comp);
  std::vector<int> mi, mj;
  for (auto [i, j] : nz) {
    mi.push_back(i);
    mj.push_back(j);
  }
  o << "constexpr int NMC = " << nz.size() << ";\n";
  o << "__device__ constexpr int MCI[NMC > 0 ? NMC : 1] = " << (mi.empty() ? std::string("{0}") : int_init(mi)) << ";\n";
  o << "__device__ constexpr int MCJ[NMC > 0 ? NMC : 1] = " << (mj.empty() ? std::string("{0}") : int_init(mj)) << ";\n";
  // unit stiffness tables: sum_q FAC grad_k phi_a grad_l phi_c per class
  const int d = g.dim;
  std::vector<double> kdu;
  for (int p = 0; p < g.nclass; ++p)
    for (int a = 0; a < T.nv; ++a)
      for (int c = 0; c < T.nv; ++c)
        for (int k = 0; k < d; ++k)
          for (int l = 0; l < d; ++l) {
            double sum = 0;
            for (int q = 0; q < T.nq; ++q)
              sum += T.fac[q] * T.gphi[((static_cast<size_t>(p) * T.nq + q) * T.nv + a) * d + k] *
                     T.gphi[((static_cast<size_t>(p) * T.nq + q) * T.nv + c) * d + l];
            kdu.push_back(sum);
          }
  o << "__device__ constexpr double KDU[NCLASS][NV][NV][DIM][DIM] = " << array_init(kdu) << ";\n";
  // all diffusion numeric: the folded per-pair table KD is used instead
  bool all_const = !geo;
  for (int s : m.comp_species[comp])
    for (const auto& dt : m.species[s].diffusion) all_const = all_const && tensor_const(dt.D);
  o << "constexpr bool CONST_DIFFUSION = " << (all_const ? "true" : "false") << ";\n";
  // nonzero entries (i, j, k, l) of the spatially uniform diffusion tensors
  // D_ij, as a compact list: dv[n] = D_{DVI[n] DVJ[n]}[DVK[n]][DVL[n]]
  auto vars = volume_vars(m, comp, T.evol, geo);
  const auto& species = m.comp_species[comp];
  std::vector<int> di, dj, dk, dl;
  std::vector<std::string> dexpr;
  for (size_t i = 0; i < species.size(); ++i)
    for (const auto& dt : m.species[species[i]].diffusion) {
      const int j = m.species[dt.wrt].local;
      for (int k = 0; k < d; ++k)
        for (int l = 0; l < d; ++l) {
          Ex e = dt.D.tensor ? dt.D.d[k][l] : (k == l ? dt.D.d[0][0] : nullptr);
          if (!e) continue;
          di.push_back(static_cast<int>(i));
          dj.push_back(j);
          dk.push_back(k);
          dl.push_back(l);
          std::string v = expr::emit(e, vars);
          // unit-cell table -> cell of spacing h: |cell| / (h_k h_l)
          if (geo) v = "(" + v + ") * prm[6] * prm[" + std::to_string(3 + k) + "] * prm[" + std::to_string(3 + l) + "]";
          dexpr.push_back(v);
        }
    }
  const int ndv = static_cast<int>(di.size());
  o << "constexpr int NDV = " << ndv << ";\n";
  auto ints = [&](const std::vector<int>& v) { return v.empty() ? std::string("{0}") : int_init(v); };
  o << "__device__ constexpr int DVI[NDV > 0 ? NDV : 1] = " << ints(di) << ";\n";
  o << "__device__ constexpr int DVJ[NDV > 0 ? NDV : 1] = " << ints(dj) << ";\n";
  o << "__device__ constexpr int DVK[NDV > 0 ? NDV : 1] = " << ints(dk) << ";\n";
  o << "__device__ constexpr int DVL[NDV > 0 ? NDV : 1] = " << ints(dl) << ";\n";
  // entries kept per block by the row builder: the point-coefficient
  // pattern and the diffusion pairs; their positions in the accumulator
  {
    std::vector<std::pair<int, int>> zb = nz;
    auto zindex = [&](int i, int j) {
      for (size_t z = 0; z < zb.size(); ++z)
        if (zb[z].first == i && zb[z].second == j) return static_cast<int>(z);
      zb.push_back({i, j});
      return static_cast<int>(zb.size() - 1);
    };
    std::vector<int> mcz, dvz, kzi, kzj, kzz;
    for (auto [i, j] : nz) mcz.push_back(zindex(i, j));
    for (size_t n = 0; n < di.size(); ++n) dvz.push_back(zindex(di[n], dj[n]));
    for (size_t n = 0; n < di.size(); ++n) {
      bool seen = false;
      for (size_t k2 = 0; k2 < kzi.size(); ++k2) seen = seen || (kzi[k2] == di[n] && kzj[k2] == dj[n]);
      if (seen) continue;
      kzi.push_back(di[n]);
      kzj.push_back(dj[n]);
      kzz.push_back(zindex(di[n], dj[n]));
    }
    const int bcn = static_cast<int>(species.size());
    std::vector<int> zix(static_cast<size_t>(bcn) * bcn, -1);
    for (size_t z = 0; z < zb.size(); ++z) zix[zb[z].first * bcn + zb[z].second] = static_cast<int>(z);
    auto arr = [&](const std::vector<int>& v) { return v.empty() ? std::string("{0}") : int_init(v); };
    o << "constexpr int NZB = " << zb.size() << ", NKZ = " << kzi.size() << ";\n";
    o << "__device__ constexpr int ZIX[BC][BC] = " << arr(zix) << ";\n";
    o << "__device__ constexpr int MCZ[NMC > 0 ? NMC : 1] = " << arr(mcz) << ";\n";
    o << "__device__ constexpr int DVZ[NDV > 0 ? NDV : 1] = " << arr(dvz) << ";\n";
    o << "__device__ constexpr int KZI[NKZ > 0 ? NKZ : 1] = " << arr(kzi) << ";\n";
    o << "__device__ constexpr int KZJ[NKZ > 0 ? NKZ : 1] = " << arr(kzj) << ";\n";
    o << "__device__ constexpr int KZZ[NKZ > 0 ? NKZ : 1] = " << arr(kzz) << ";\n";
  }
  o << "__device__ __forceinline__ void uniform_diffusion(double t, const double* __restrict__ prm, "
       "double (&dv)[NDV > 0 ? NDV : 1]) {\n  (void)t; (void)prm; dv[0] = 0.0;\n";
  for (int n = 0; n < ndv; ++n) o << "  dv[" << n << "] = " << dexpr[n] << ";\n";
  o << "}\n";
  o << "}  // namespace " << ns << "\n";
  std::string body = kTiled;
  if (points) body += kTiledPoints;
  if (geo && comp < 4) body += kTiledBatch;
  if (geo && comp < 4 && points) body += kTiledPointBatch;
  {
    // one point apply (and batched point apply) per candidate tile shape k:
    // %KERNEL%_p<k>, _pf<k>, _pbatch<k>, _pfbatch<k>
    std::string pe, pb;
    for (size_t k2 = 0; k2 < apply_tiles.size(); ++k2) {
      const std::string id = std::to_string(k2);
      const std::string tl = ns + "::ApplyShape" + id;
      pe += "DCG_POINT_ENTRY(%KERNEL%_p" + id + ", double, " + tl + ")\n";
      pe += "DCG_POINT_ENTRY(%KERNEL%_pf" + id + ", float, " + tl + ")\n";
      pb += "DCG_PBATCH_ENTRY(%KERNEL%_pbatch" + id + ", double, v, pc, " + tl + ")\n";
      pb += "DCG_PBATCH_ENTRY(%KERNEL%_pfbatch" + id + ", float, f, pc32, " + tl + ")\n";
    }
    // the candidate apply shapes as plain names (macro arguments hold no commas)
    std::string as;
    for (size_t k2 = 0; k2 < apply_tiles.size(); ++k2)
      as += "using ApplyShape" + std::to_string(k2) + " = ApplyTile<" + std::to_string(apply_tiles[k2][0]) + ", " +
            std::to_string(apply_tiles[k2][1]) + ", " + std::to_string(apply_tiles[k2][2]) + ", " +
            std::to_string(apply_tile_threads(g.dim, apply_tiles[k2], static_cast<int>(m.comp_species[comp].size()))) +
            ", " + std::to_string(k2 < apply_waves.size() ? apply_waves[k2] : 1) + ">;\n";
    replace_all(body, "%APPLY_SHAPES%", as);
    replace_all(body, "%POINT_ENTRIES%", pe);
    replace_all(body, "%PBATCH_ENTRIES%", pb);
  }
  replace_all(body, "%C%", ns);
  replace_all(body, "%KERNEL%", kname);
  replace_all(body, "%COMP%", std::to_string(comp));
  return o.str() + body;
}

GeneratedSource generate_assembly(const ModelSpec& spec, const Layout& L, bool runtime_geometry, bool points) {
  GeneratedSource out;
  const bool geo = runtime_geometry;
  std::string pre = kPrelude;
  replace_all(pre, "%B%", std::to_string(L.B));
  replace_all(pre, "%GEO%", geo ? "1" : "0");
  replace_all(pre, "%UNS%", L.grid.unstructured ? "1" : "0");
  replace_all(pre, "%ASM_BYTES%", std::to_string(sizeof(k::AsmMember)));
  replace_all(pre, "%PCM_BYTES%", std::to_string(sizeof(k::PcMember)));
  std::ostringstream src;
  src << pre;
  const bool numerical = spec.jacobian == JacobianType::Numerical && !spec.is_linear;
  for (size_t c = 0; c < spec.comp_names.size(); ++c) {
    if (spec.comp_species[c].empty()) {
      out.volume_kernel.push_back("");
      out.tile_kernel.push_back("");
      out.apply_tiles.push_back({});
      out.apply_waves.push_back({});
      out.tile_batch_kernel.push_back("");
      out.tile_dims.push_back({0, 0, 0});
      out.point_words.push_back(0);
      continue;
    }
    std::string ns = "vol" + std::to_string(c);
    std::string kn = "dcg_volume_" + std::to_string(c);
    const bool tiled = !L.grid.unstructured && tile_eligible(spec, static_cast<int>(c));
    src << gen_volume(spec, L.grid, static_cast<int>(c), ns, kn, numerical, geo, !tiled);
    out.volume_kernel.push_back(tiled ? "" : kn);
    if (tiled) {
      auto tile = choose_tile(L.grid, static_cast<int>(spec.comp_species[c].size()),
                              static_cast<int>(mc_pattern(spec, static_cast<int>(c)).size()));
      std::string tk = "dcg_tile_" + std::to_string(c);
      std::vector<int> waves;
      const auto shapes = points ? apply_tile_candidates(L.grid, static_cast<int>(spec.comp_species[c].size()), tile,
                                                         &waves)
                                 : std::vector<std::array<int, 3>>{};
      src << gen_tiled(spec, L.grid, static_cast<int>(c), ns, tk, tile, geo, points, shapes, waves);
      out.tile_kernel.push_back(tk);
      out.apply_tiles.push_back(shapes);
      out.apply_waves.push_back(waves);
      out.tile_batch_kernel.push_back(geo && c < 4 ? tk + "_batch" : "");
      out.tile_dims.push_back(tile);
      const int nq = static_cast<int>(element_rule(L.grid.dim, L.grid.cube).w.size());
      out.point_words.push_back(points ? L.grid.nclass * nq *
                                             static_cast<int>(mc_pattern(spec, static_cast<int>(c)).size())
                                       : 0);
    } else {
      out.tile_kernel.push_back("");
      out.apply_tiles.push_back({});
      out.apply_waves.push_back({});
      out.tile_batch_kernel.push_back("");
      out.tile_dims.push_back({0, 0, 0});
      out.point_words.push_back(0);
    }
  }
  // facet kernels for every (in, out, boundary) combination present
  std::vector<std::tuple<int, int, bool>> combos;
  for (const auto& F : L.facets) {
    auto key = std::make_tuple(F.comp_in, F.comp_out, F.boundary);
    bool seen = false;
    for (auto& k : combos) seen = seen || k == key;
    if (!seen) combos.push_back(key);
  }
  for (auto& [ci, co, bnd] : combos) {
    FacetKernel fk;
    fk.comp_in = ci;
    fk.comp_out = co;
    fk.boundary = bnd;
    fk.name = "dcg_facet_" + std::to_string(ci) + "_" + std::to_string(co) + (bnd ? "_b" : "_i");
    std::string ns = "fac" + std::to_string(ci) + "_" + std::to_string(co) + (bnd ? "b" : "i");
    src << gen_facet(spec, L.grid, ci, co, bnd, ns, fk.name, numerical, fk);
    out.facet_kernels.push_back(fk);
  }
  out.source = src.str();
  return out;
}

double cell_det(const Grid& g, long cell) {
  const int d = g.dim;
  double x0[3], xk[3], J[3][3] = {{0}}, Ji[3][3] = {{0}}, det = 0;
  g.vertex_pos(g.cell_vertex(cell, 0), x0);
  for (int col = 0; col < d; ++col) {
    g.vertex_pos(g.cell_vertex(cell, col + 1), xk);
    for (int k = 0; k < d; ++k) J[k][col] = xk[k] - x0[k];
  }
  invert(d, J, Ji, det);
  return det;
}

std::vector<double> cell_geometry(const Grid& g, const std::vector<int>& cells, const Rule* rule) {
  const int d = g.dim, nv = g.nvpe;
  const Rule r = rule ? *rule : element_rule(d, false);
  const int nq = static_cast<int>(r.w.size());
  std::vector<double> out;
  out.reserve(cells.size() * (nv * d + nq * 3 + 1));
  double phi[8], dref[8][3];
  const double xi0[3] = {0, 0, 0};
  shape(d, false, xi0, phi, dref);  // P1: constant reference gradients
  for (int cell : cells) {
    double x0[3], xk[3], J[3][3] = {{0}}, Ji[3][3] = {{0}}, det = 0;
    g.vertex_pos(g.cell_vertex(cell, 0), x0);
    for (int col = 0; col < d; ++col) {
      g.vertex_pos(g.cell_vertex(cell, col + 1), xk);
      for (int k = 0; k < d; ++k) J[k][col] = xk[k] - x0[k];
    }
    invert(d, J, Ji, det);
    if (!(det > 0.0)) throw Error("degenerate cell " + std::to_string(cell) + " in the grid file");
    for (int a = 0; a < nv; ++a)
      for (int k = 0; k < d; ++k) {
        double s = 0;
        for (int l = 0; l < d; ++l) s += dref[a][l] * Ji[l][k];
        out.push_back(s);
      }
    for (int q = 0; q < nq; ++q)
      for (int k = 0; k < 3; ++k) {
        double s = k < d ? x0[k] : 0.0;
        for (int l = 0; l < d && k < d; ++l) s += J[k][l] * r.x[q][l];
        out.push_back(s);
      }
    out.push_back(det);
  }
  return out;
}

std::string generate_reduce(const ModelSpec& m, const Layout& L, const std::vector<ReduceKey>& keys,
                            std::vector<std::string>* kernels) {
  const Grid& g = L.grid;
  const Rule rule = reduce_rule(g.dim, g.cube);
  const auto T = element_tables(unit_cell(g), &rule);
  const int nk = static_cast<int>(keys.size());
  std::ostringstream src;
  src << "#define DCG_B " << L.B << "\n";
  kernels->clear();
  for (size_t c = 0; c < L.comps.size(); ++c) {
    const auto& C = L.comps[c];
    if (C.cells.empty() || nk == 0) {
      kernels->push_back("");
      continue;
    }
    const int bc = C.b;
    const std::string ns = "red" + std::to_string(c), kn = "dcg_reduce_" + std::to_string(c);
    auto vars = volume_vars(m, static_cast<int>(c), T.evol, true);
    // reduce binds position, time and integration_factor only: the other
    // point keywords keep LocalDomain's zero defaults (local_domain.hh:40-46)
    vars["integration_factor"] = "ifac";
    vars["entity_volume"] = "0.0";
    vars["in_volume"] = "0.0";
    std::ostringstream o;
    o << "namespace " << ns << " {\n";
    o << "constexpr int DIM = " << g.dim << ", NV = " << T.nv << ", NQR = " << T.nq << ", NCLASS = " << g.nclass
      << ", NCORNER = " << g.ncorner << ", BC = " << (bc > 0 ? bc : 1) << ", NK = " << nk << ";\n";
    // cells of grids read from files carry their geometry (cell_geometry with this rule)
    o << "constexpr bool UNS = " << (g.unstructured ? "true" : "false")
      << ";\nconstexpr int RG_POS = NV * DIM, RG_DET = RG_POS + NQR * 3, RG_WORDS = RG_DET + 1;\n";
    std::vector<int> corner;
    for (int p = 0; p < g.nclass; ++p)
      for (int a = 0; a < 8; ++a) corner.push_back(g.corner[p][a]);
    o << "__device__ constexpr int CORNER[NCLASS][8] = " << int_init(corner) << ";\n";
    o << "__device__ constexpr double RPHI[NQR][NV] = " << array_init(T.phi) << ";\n";
    o << "__device__ constexpr double RG[NCLASS][NQR][NV][DIM] = " << array_init(T.gphi) << ";\n";
    o << "__device__ constexpr double RX[NCLASS][NQR][3] = " << array_init(T.xq) << ";\n";
    o << "__device__ constexpr double RW[NQR] = " << array_init(T.fac) << ";\n";
    // reductions (point value, accumulator) -> accumulator
    for (int k = 0; k < nk; ++k) {
      const auto& f = keys[k].reduction;
      o << "__device__ __forceinline__ double red" << k << "(double a, double b, const double* __restrict__ prm) {\n"
        << "  (void)prm;\n";
      if (f.body) {
        auto rv = vars;
        rv[f.args[0]] = "a";
        rv[f.args[1]] = "b";
        o << "  return " << expr::emit(f.body, rv) << ";\n";
      } else if (!f.args.empty()) {
        o << "  return 0.0;\n";  // a literal zero body
      } else {
        o << "  return a + b;\n";
      }
      o << "}\n";
    }
    o << "__device__ __forceinline__ void evaluate(const double (&u)[BC], const double (&g)[BC][3], double px, "
         "double py, double pz, double t, double ifac, double hv, const double* __restrict__ prm, double (&e)[NK]) {\n"
         "  (void)u; (void)g; (void)px; (void)py; (void)pz; (void)t; (void)ifac; (void)hv; (void)prm;\n";
    for (int k = 0; k < nk; ++k)
      o << "  e[" << k << "] = " << (keys[k].evaluation ? expr::emit(keys[k].evaluation, vars) : std::string("0.0"))
        << ";\n";
    o << "}\n";
    o << "__device__ __forceinline__ double combine(int k, double a, double b, const double* __restrict__ prm) {\n"
         "  switch (k) {\n";
    for (int k = 0; k < nk; ++k) o << "    case " << k << ": return red" << k << "(a, b, prm);\n";
    o << "  }\n  return a + b;\n}\n";
    o << "}  // namespace " << ns << "\n";
    std::string body = R"(
extern "C" __global__ void __launch_bounds__(256)
%KERNEL%(const int* __restrict__ cells, const unsigned char* __restrict__ cmask, const int* __restrict__ cslots,
         int ncell, const double* __restrict__ x, double t, int n0, int n1, double ox, double oy, double oz,
         double hx, double hy, double hz, const double* __restrict__ prm, double* __restrict__ part,
         const double* __restrict__ rgeo) {
  using namespace %C%;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double h[3] = {hx, hy, hz};
  const double hv = hx * (DIM > 1 ? hy : 1.0) * (DIM > 2 ? hz : 1.0);
  double acc[NK];
  bool have[NK];
#pragma unroll
  for (int k = 0; k < NK; ++k) {
    acc[k] = 0.0;
    have[k] = false;
  }
  if (i < ncell) {
    const int cell = cells[i];
    const int ci = cell % n0, cj = (cell / n0) % n1, ck = cell / (n0 * n1);
    const double cx = ox + ci * hx, cy = DIM > 1 ? oy + cj * hy : 0.0, cz = DIM > 2 ? oz + ck * hz : 0.0;
    const double* eg = UNS ? rgeo + (size_t)i * RG_WORDS : nullptr;
    const unsigned m = cmask[i];
    for (int p = 0; p < NCLASS; ++p) {
      if (!((m >> p) & 1u)) continue;
      double ul[NV][BC];
      for (int a = 0; a < NV; ++a) {
        const int s = cslots[(size_t)i * NCORNER + CORNER[p][a]];
        for (int b = 0; b < BC; ++b) ul[a][b] = (s >= 0 && b < %NB%) ? x[(size_t)s * DCG_B + b] : 0.0;
      }
      for (int q = 0; q < NQR; ++q) {
        double u[BC], gr[BC][3];
        for (int b = 0; b < BC; ++b) {
          u[b] = 0.0;
          gr[b][0] = gr[b][1] = gr[b][2] = 0.0;
          for (int a = 0; a < NV; ++a) {
            u[b] += RPHI[q][a] * ul[a][b];
            for (int k = 0; k < DIM; ++k)
              gr[b][k] += (UNS ? eg[a * DIM + k] : RG[p][q][a][k] / h[k]) * ul[a][b];
          }
        }
        double px, py, pz, ifac;
        if constexpr (UNS) {
          px = eg[RG_POS + q * 3];
          py = DIM > 1 ? eg[RG_POS + q * 3 + 1] : 0.0;
          pz = DIM > 2 ? eg[RG_POS + q * 3 + 2] : 0.0;
          ifac = RW[q] * eg[RG_DET];
        } else {
          px = cx + RX[p][q][0] * hx;
          py = DIM > 1 ? cy + RX[p][q][1] * hy : 0.0;
          pz = DIM > 2 ? cz + RX[p][q][2] * hz : 0.0;
          ifac = RW[q] * hv;
        }
        double e[NK];
        evaluate(u, gr, px, py, pz, t, ifac, hv, prm, e);
#pragma unroll
        for (int k = 0; k < NK; ++k) {
          acc[k] = have[k] ? combine(k, e[k], acc[k], prm) : e[k];
          have[k] = true;
        }
      }
    }
  }
  // threads in a fixed tree: the higher half is the later cells
  __shared__ double sv[256];
  __shared__ int sh[256];
  for (int k = 0; k < NK; ++k) {
    sv[threadIdx.x] = acc[k];
    sh[threadIdx.x] = have[k] ? 1 : 0;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
      if (threadIdx.x < st) {
        const int o = threadIdx.x + st;
        if (sh[o]) {
          sv[threadIdx.x] = sh[threadIdx.x] ? combine(k, sv[o], sv[threadIdx.x], prm) : sv[o];
          sh[threadIdx.x] = 1;
        }
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      part[((size_t)blockIdx.x * NK + k) * 2] = sv[0];
      part[((size_t)blockIdx.x * NK + k) * 2 + 1] = sh[0];
    }
    __syncthreads();
  }
}
)";
    replace_all(body, "%C%", ns);
    replace_all(body, "%KERNEL%", kn);
    replace_all(body, "%NB%", std::to_string(bc));
    src << o.str() << body;
    kernels->push_back(kn);
  }
  return src.str();
}

}  // namespace dcopasi
