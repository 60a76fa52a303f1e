// Transform/reduce over the mesh: the reference's `model.reduce` keys
// (dune/copasi/model/diffusion_reaction/reduce.hh:40-285, called by the CLI
// after every step, src/dune_copasi.cc:420-423).  For each key,
//   acc = initial.value; for every quadrature point (order 4) of every cell:
//     acc = reduction(evaluation, acc)          (default: +)
//   value = transformation(reduction(acc, initial.value))
// then `error.expression(value) != 0` raises a reduction error and
// `warn.expression(value) != 0` a warning.  Here the evaluations run in one
// generated kernel per compartment (hipRTC) that folds each thread's points
// and combines the threads in a fixed tree; the host folds the workgroup
// partials in order and applies initial value, transformation, error and
// warn.  For associative reductions (sum, max, min, product) this equals the
// sequential order up to rounding.
#pragma once
#include <map>
#include <string>
#include <vector>

#include "config.hh"
#include "expr.hh"

namespace dcopasi {

struct ReduceFunction {
  std::vector<std::string> args;
  expr::Ex body;  // null: absent
};

struct ReduceKey {
  std::string name;
  expr::Ex evaluation;       // null: 0
  ReduceFunction reduction;  // two arguments (point value, accumulator); absent: +
  double initial = 0.0;
  ReduceFunction transformation, error, warn;  // one argument each; absent: none
  bool quiet = false;
};

struct ReduceResult {
  std::string name;
  double value = 0.0;
  int flag = 0;  // 0, 1 warn expression true, 2 error expression true
  std::string message;
};

// keys under `prefix` (e.g. "model.reduce") in first-appearance order
std::vector<ReduceKey> read_reduce(const ParameterTree& cfg, const std::string& prefix, const expr::Context& ctx);

// a(b...) of a parsed "args: body" function on the host
double reduce_call(const ReduceFunction& f, const std::vector<double>& x, const std::map<std::string, double>& params);

}  // namespace dcopasi
