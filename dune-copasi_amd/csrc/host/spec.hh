// Model description read from the INI keys of the reference's diffusion-
// reaction model (model/diffusion_reaction/local_equations.hh:263-355,551-700,
// functor_factory_parser.impl.hh:24-112):
//   model.scalar_field.<s>.compartment
//   model.scalar_field.<s>.storage.expression            (+ .jacobian.<w>.expression)
//   model.scalar_field.<s>.reaction.expression           (+ .jacobian.<w>.expression)
//   model.scalar_field.<s>.velocity.<x|y|z>.expression   (+ .jacobian.<w>.<x|y|z>.expression)
//   model.scalar_field.<s>.cross_diffusion.<w>[.type=scalar|tensor]
//        .expression | .<xx|xy|...>.expression           (+ .jacobian.<v>...)
//   model.scalar_field.<s>.outflow.<compartment>.expression (+ .jacobian.<w>.expression)
// Storage terms form the mass operator, the rest the stiffness operator
// (local_equations.hh:322-355).
#pragma once
#include <array>
#include <map>
#include <string>
#include <vector>

#include "config.hh"
#include "expr.hh"

namespace dcopasi {

using expr::Ex;

struct Partial {
  int wrt;  // global species index
  Ex e;
};

struct TensorTerm {
  bool tensor = false;
  Ex d[3][3];  // scalar diffusion in d[0][0]
};

struct DiffusionTerm {
  int wrt;  // global species index (same compartment)
  TensorTerm D;
  std::vector<std::pair<int, TensorTerm>> jac;
};

struct OutflowTerm {
  int to_comp;
  Ex e;
  std::vector<Partial> jac;
};

struct SpeciesSpec {
  std::string name;
  int comp = -1, local = -1;
  Ex storage, reaction;
  std::vector<Partial> storage_jac, reaction_jac;
  std::array<Ex, 3> velocity{};
  std::vector<std::pair<int, std::array<Ex, 3>>> velocity_jac;
  std::vector<DiffusionTerm> diffusion;
  std::vector<OutflowTerm> outflow;
};

enum class JacobianType { Analytical, Numerical, Symbolic };

struct ModelSpec {
  int dim = 2;
  std::vector<std::string> comp_names;
  std::vector<std::vector<int>> comp_species;
  std::vector<Ex> comp_expr;
  std::vector<SpeciesSpec> species;
  std::vector<Ex> initial;  // per species, may be null
  expr::Context ctx;
  JacobianType jacobian = JacobianType::Analytical;
  double fd_epsilon = 1e-7;
  bool is_linear = false;
  bool has_outflow_boundary = false, has_outflow_interface = false;
  // run-time parameters (parser_context constants when not folded): the
  // expressions reference "$<name>", constants sorted by name
  std::vector<std::string> params;
  std::vector<double> param_values;
};

// fold_constants = false keeps parser_context constants as run-time
// parameters so models that differ only in them share generated code.
ModelSpec read_model(const ParameterTree& cfg, int dim, bool fold_constants = true);

}  // namespace dcopasi
