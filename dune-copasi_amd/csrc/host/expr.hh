// Math expressions of DuneCopasi model files, compiled instead of interpreted.
//
// The reference evaluates each `*.expression` key with a runtime parser at
// every quadrature point (binding in
// dune/copasi/model/functor_factory_parser.impl.hh:114-182, language in
// doc/docusaurus/docs/math_expr.mdx).  Here an expression is parsed once into
// an immutable tree, parser_context constants are folded and parser_context
// functions ("args: body") are inlined, and the tree is either
//   * emitted as C++ for the device assembly kernels (compiled with hipRTC),
//   * differentiated symbolically (for Jacobian terms a model leaves out), or
//   * evaluated on the host (initial conditions, compartment marking).
// '^' binds tighter than unary minus and associates to the right; the
// constant `no_value` is DBL_MAX.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace dcopasi::expr {

enum class Kind { Const, Var, Neg, Not, Add, Sub, Mul, Div, Pow, Mod, Lt, Le, Gt, Ge, Eq, Ne, And, Or, Cond, Call };

struct Node;
using Ex = std::shared_ptr<const Node>;

struct Node {
  Kind kind;
  double value = 0.0;  // Const
  std::string name;    // Var: variable name, Call: function name
  std::vector<Ex> args;
};

struct Function {
  std::vector<std::string> params;
  std::string body;
};

// Symbols visible to expressions of one model.
struct Context {
  std::map<std::string, double> constants;
  std::map<std::string, Function> functions;
  // false: constants stay symbolic as variables named "$<name>" (model
  // parameters read at run time, shared code for a batch of models)
  bool fold_constants = true;
};

// Parses `text`; variables are left symbolic, context constants fold to
// numbers and context functions are expanded.  An empty string or a literal
// zero ("0", "0.0", "-0") yields nullptr: the term is absent, as in
// functor_factory_parser.impl.hh:119-122.
Ex parse(const std::string& text, const Context& ctx);

// Constructors with constant folding and the usual identities.
Ex constant(double v);
Ex variable(const std::string& name);
Ex make(Kind k, std::vector<Ex> args);
Ex call(const std::string& fn, std::vector<Ex> args);

bool is_const(const Ex& e, double* v = nullptr);
bool depends_on(const Ex& e, const std::string& var);
void collect_vars(const Ex& e, std::vector<std::string>& out);

// d e / d var; nullptr when identically zero.
Ex derivative(const Ex& e, const std::string& var);

// C++ source for the device; `map` renames variables to C expressions.
std::string emit(const Ex& e, const std::map<std::string, std::string>& map);

// Host evaluation with variables looked up in `vars` (missing -> error).
double evaluate(const Ex& e, const std::map<std::string, double>& vars);

// Bind variable names to slots once, then evaluate repeatedly.
struct Bound {
  std::vector<int> code;  // compact postfix program
  std::vector<double> consts;
  double operator()(const double* slots) const;
};
Bound bind(const Ex& e, const std::map<std::string, int>& slots);

}  // namespace dcopasi::expr
