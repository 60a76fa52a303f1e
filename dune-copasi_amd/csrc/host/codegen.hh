// Generates the HIP source of a model's assembly kernels (compiled once per
// model with hipRTC).  One volume kernel per compartment restates
// LocalOperator::localAssembleVolume / localAssembleAnalyticalJacobianVolume /
// localAssembleNumericalJacobianVolume (local_operator.hh:417-491,541-765)
// with the model's expressions inlined; one facet kernel per compartment pair
// (and per compartment for domain-boundary outflow) restates
// localAssembleSkeleton / localAssembleBoundary and their Jacobians
// (local_operator.hh:777-1452) with exact derivatives in all four blocks.
#pragma once
#include <array>
#include <string>
#include <vector>

#include "layout.hh"
#include "spec.hh"

namespace dcopasi {

struct FacetKernel {
  int comp_in, comp_out;
  bool boundary;
  std::string name;
  int nvf, nqf, fint, fdbl;  // vertices, quadrature points, ints and doubles per facet record
};

struct GeneratedSource {
  std::string source;
  // coloured kernel per compartment ("" when it has no species or is tiled)
  std::vector<std::string> volume_kernel;
  // tiled write-once variant per compartment ("" when the model needs the coloured kernel)
  std::vector<std::string> tile_kernel;
  // batched tiled kernel per compartment (runtime geometry only, else "")
  std::vector<std::string> tile_batch_kernel;
  std::vector<std::array<int, 3>> tile_dims;
  // words per cell of the point-coefficient Jacobian (classes x points x
  // coefficients; 0 when the compartment has no point kernels), stored for
  // every cell of the grid, word-major: pc[word * cells + cell]
  std::vector<int> point_words;
  // candidate tile shapes of the point apply per compartment (entry points
  // <tile_kernel>_p<k>, _pf<k> and the batched _pbatch<k>, _pfbatch<k>);
  // the first is the assembly tile
  std::vector<std::vector<std::array<int, 3>>> apply_tiles;
  // per compartment and candidate: the waves per SIMD its registers are bounded for (1: none)
  std::vector<std::vector<int>> apply_waves;
  std::vector<FacetKernel> facet_kernels;
};

// Words of the run-time parameter array before the model parameters: the
// cell spacing, its reciprocals and the cell measure (read by kernels generated
// with runtime_geometry, zero-cost constants otherwise).
constexpr int kGeoWords = 8;

// runtime_geometry: element tables of the unit cell, spacing from prm[0..7],
// so models that differ only in mesh size and parameters share one source.
// the compartment's equations suit the tiled kernel (analytic Jacobian, no
// velocity, diffusion uniform in space and independent of the state)
bool tile_eligible(const ModelSpec& m, int comp);
// tile shapes the point apply is generated for (the set-up tile first)
std::vector<std::array<int, 3>> apply_tile_candidates(const Grid& g, int species, std::array<int, 3> setup_tile,
                                                      std::vector<int>* waves = nullptr);
// workgroup size of the point apply with tile t (cells per workgroup along x, y, z)
int apply_tile_threads(int dim, std::array<int, 3> t, int species);

// points: also emit the point-coefficient Jacobian entry points (_q, _p, _pf)
// of every tiled compartment
GeneratedSource generate_assembly(const ModelSpec& spec, const Layout& layout, bool runtime_geometry = false,
                                  bool points = false);

// Geometry words of unstructured cells, in the order of `cells`, as the
// generated cell kernels read them: the gradients of the element's basis
// functions [a][k], the points of the order-2 rule [q][3], |det J|.
// `rule`: the points to place (default: the order-2 rule of the assembly).
std::vector<double> cell_geometry(const Grid& g, const std::vector<int>& cells, const Rule* rule = nullptr);
// |det J| of one unstructured cell
double cell_det(const Grid& g, long cell);

struct ReduceKey;
// Device source of a transform/reduce evaluation: kernel "dcg_reduce_<c>" per
// compartment with cells ("" otherwise) folds every key's evaluation over the
// order-4 points of its cells, one thread per cell, and writes one partial
// (value, defined) per key and workgroup.  Unit-cell tables; the spacing,
// origin and time come at run time.
std::string generate_reduce(const ModelSpec& spec, const Layout& layout, const std::vector<ReduceKey>& keys,
                            std::vector<std::string>* kernels);

}  // namespace dcopasi
