// Oracle math-expression interpreter.
//
// Restates the expression language DuneCopasi evaluates at every quadrature
// point through its runtime parsers (dune/copasi/parser/*.hh; variable binding
// in dune/copasi/model/functor_factory_parser.impl.hh:133-169; language in
// doc/docusaurus/docs/math_expr.mdx): + - * / ^, unary minus, comparisons,
// and/or/not, the ternary "c ? a : b", the usual elementary functions, the
// constant `no_value` (DBL_MAX), and parser_context constants and inline
// functions "args: body" (src/dune/copasi/parser/context.cc).
//
// The oracle evaluates the tree directly, as the reference's ExprTk/muParser
// interpreters do; it shares no code with the product's compiler.
// Operator precedence: '^' binds tighter than unary minus (-x^2 == -(x^2)) and
// is right associative.
//
// Test infrastructure only.
#pragma once
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ini.hh"

namespace orc {

enum class Op {
  Num, Var, Neg, Not, Add, Sub, Mul, Div, Pow, Mod,
  Lt, Le, Gt, Ge, Eq, Ne, And, Or, Tern, Call, Table
};

// parser_context type "interpolation": points (domain sorted) of a piecewise
// linear function, constant beyond the ends (context.cc:72-97)
struct Table {
  std::vector<double> x, y;
};

struct Node {
  Op op;
  double num = 0;
  int var = -1;
  std::string fname;
  std::shared_ptr<const Table> table;  // Op::Table
  std::vector<std::shared_ptr<Node>> kids;
};
using NodeP = std::shared_ptr<Node>;

struct UserFunction {
  std::vector<std::string> args;
  std::string body;
};

// Name resolution for a parse: variables map to slots, constants fold in,
// user functions are expanded inline at their call sites.
struct Scope {
  std::map<std::string, int> vars;
  std::map<std::string, double> consts;
  std::map<std::string, UserFunction> funcs;
  std::map<std::string, std::shared_ptr<const Table>> tables;
};

class Parser {
 public:
  Parser(const std::string& src, const Scope& scope,
         const std::map<std::string, NodeP>* args = nullptr, int depth = 0)
      : s_(src), scope_(scope), args_(args), depth_(depth) {
    if (depth_ > 32) throw std::runtime_error("expr: function nesting too deep");
  }
  NodeP parse() {
    next();
    NodeP n = ternary();
    if (tok_ != T::End) fail("unexpected trailing input");
    return n;
  }

 private:
  enum class T { End, Num, Id, Op, LPar, RPar, Comma, Quest, Colon };
  const std::string s_;
  const Scope& scope_;
  const std::map<std::string, NodeP>* args_;
  int depth_;
  size_t pos_ = 0;
  T tok_ = T::End;
  std::string text_;
  double val_ = 0;

  [[noreturn]] void fail(const std::string& m) {
    throw std::runtime_error("expr: " + m + " in '" + s_ + "'");
  }
  void next() {
    while (pos_ < s_.size() && isspace((unsigned char)s_[pos_])) ++pos_;
    if (pos_ >= s_.size()) { tok_ = T::End; return; }
    char c = s_[pos_];
    if (isdigit((unsigned char)c) || (c == '.' && pos_ + 1 < s_.size() && isdigit((unsigned char)s_[pos_ + 1]))) {
      size_t used = 0;
      val_ = std::stod(s_.substr(pos_), &used);
      pos_ += used;
      tok_ = T::Num;
      return;
    }
    if (isalpha((unsigned char)c) || c == '_') {
      size_t b = pos_;
      while (pos_ < s_.size() && (isalnum((unsigned char)s_[pos_]) || s_[pos_] == '_')) ++pos_;
      text_ = s_.substr(b, pos_ - b);
      if (text_ == "and" || text_ == "or" || text_ == "not") { tok_ = T::Op; return; }
      tok_ = T::Id;
      return;
    }
    ++pos_;
    switch (c) {
      case '(': tok_ = T::LPar; return;
      case ')': tok_ = T::RPar; return;
      case ',': tok_ = T::Comma; return;
      case '?': tok_ = T::Quest; return;
      case ':': tok_ = T::Colon; return;
      default: break;
    }
    text_ = std::string(1, c);
    if (pos_ < s_.size()) {
      std::string two = text_ + s_[pos_];
      if (two == "<=" || two == ">=" || two == "==" || two == "!=" || two == "&&" || two == "||") {
        text_ = two;
        ++pos_;
      }
    }
    if (text_ == "&&") text_ = "and";
    if (text_ == "||") text_ = "or";
    if (text_ == "!") text_ = "not";
    static const char* ok[] = {"+", "-", "*", "/", "^", "%", "<", ">", "<=", ">=", "==", "!=", "and", "or", "not"};
    for (auto o : ok)
      if (text_ == o) { tok_ = T::Op; return; }
    fail("unknown character '" + text_ + "'");
  }
  bool is_op(const char* o) const { return tok_ == T::Op && text_ == o; }
  static NodeP mk(Op op, std::vector<NodeP> k) {
    auto n = std::make_shared<Node>();
    n->op = op;
    n->kids = std::move(k);
    return n;
  }
  static NodeP num(double v) {
    auto n = std::make_shared<Node>();
    n->op = Op::Num;
    n->num = v;
    return n;
  }
  NodeP ternary() {
    NodeP c = lor();
    if (tok_ == T::Quest) {
      next();
      NodeP a = ternary();
      if (tok_ != T::Colon) fail("expected ':'");
      next();
      NodeP b = ternary();
      return mk(Op::Tern, {c, a, b});
    }
    return c;
  }
  NodeP lor() {
    NodeP l = land();
    while (is_op("or")) { next(); l = mk(Op::Or, {l, land()}); }
    return l;
  }
  NodeP land() {
    NodeP l = equality();
    while (is_op("and")) { next(); l = mk(Op::And, {l, equality()}); }
    return l;
  }
  NodeP equality() {
    NodeP l = relation();
    while (is_op("==") || is_op("!=")) {
      Op o = is_op("==") ? Op::Eq : Op::Ne;
      next();
      l = mk(o, {l, relation()});
    }
    return l;
  }
  NodeP relation() {
    NodeP l = additive();
    while (is_op("<") || is_op("<=") || is_op(">") || is_op(">=")) {
      Op o = is_op("<") ? Op::Lt : is_op("<=") ? Op::Le : is_op(">") ? Op::Gt : Op::Ge;
      next();
      l = mk(o, {l, additive()});
    }
    return l;
  }
  NodeP additive() {
    NodeP l = multiplicative();
    while (is_op("+") || is_op("-")) {
      Op o = is_op("+") ? Op::Add : Op::Sub;
      next();
      l = mk(o, {l, multiplicative()});
    }
    return l;
  }
  NodeP multiplicative() {
    NodeP l = unary();
    while (is_op("*") || is_op("/") || is_op("%")) {
      Op o = is_op("*") ? Op::Mul : is_op("/") ? Op::Div : Op::Mod;
      next();
      l = mk(o, {l, unary()});
    }
    return l;
  }
  NodeP unary() {
    if (is_op("-")) { next(); return mk(Op::Neg, {unary()}); }
    if (is_op("+")) { next(); return unary(); }
    if (is_op("not")) { next(); return mk(Op::Not, {unary()}); }
    return power();
  }
  NodeP power() {
    NodeP b = primary();
    if (is_op("^")) {
      next();
      return mk(Op::Pow, {b, unary()});  // right associative, exponent may be signed
    }
    return b;
  }
  NodeP primary() {
    if (tok_ == T::Num) { double v = val_; next(); return num(v); }
    if (tok_ == T::LPar) {
      next();
      NodeP n = ternary();
      if (tok_ != T::RPar) fail("expected ')'");
      next();
      return n;
    }
    if (tok_ != T::Id) fail("expected operand");
    std::string id = text_;
    next();
    if (tok_ == T::LPar) {
      next();
      std::vector<NodeP> a;
      if (tok_ != T::RPar) {
        for (;;) {
          a.push_back(ternary());
          if (tok_ == T::Comma) { next(); continue; }
          break;
        }
      }
      if (tok_ != T::RPar) fail("expected ')' after arguments");
      next();
      return call(id, std::move(a));
    }
    if (args_) {
      auto it = args_->find(id);
      if (it != args_->end()) return it->second;
    }
    if (id == "no_value") return num(DBL_MAX);
    auto c = scope_.consts.find(id);
    if (c != scope_.consts.end()) return num(c->second);
    auto v = scope_.vars.find(id);
    if (v != scope_.vars.end()) {
      auto n = std::make_shared<Node>();
      n->op = Op::Var;
      n->var = v->second;
      return n;
    }
    fail("unknown symbol '" + id + "'");
  }
  NodeP call(const std::string& f, std::vector<NodeP> a) {
    auto tb = scope_.tables.find(f);
    if (tb != scope_.tables.end()) {
      if (a.size() != 1) fail("wrong argument count for '" + f + "'");
      auto n = mk(Op::Table, std::move(a));
      n->table = tb->second;
      return n;
    }
    auto uf = scope_.funcs.find(f);
    if (uf != scope_.funcs.end()) {
      if (uf->second.args.size() != a.size()) fail("wrong argument count for '" + f + "'");
      std::map<std::string, NodeP> bind;
      for (size_t i = 0; i < a.size(); ++i) bind[uf->second.args[i]] = a[i];
      Parser sub(uf->second.body, scope_, &bind, depth_ + 1);
      return sub.parse();
    }
    static const std::map<std::string, int> arity = {
        {"sin", 1}, {"cos", 1}, {"tan", 1}, {"asin", 1}, {"acos", 1}, {"atan", 1},
        {"sinh", 1}, {"cosh", 1}, {"tanh", 1}, {"exp", 1}, {"log", 1}, {"ln", 1},
        {"log10", 1}, {"log2", 1}, {"sqrt", 1}, {"abs", 1}, {"floor", 1}, {"ceil", 1},
        {"round", 1}, {"sign", 1}, {"sgn", 1}, {"pow", 2}, {"atan2", 2}, {"if", 3},
        {"min", -1}, {"max", -1}};
    auto it = arity.find(f);
    if (it == arity.end()) fail("unknown function '" + f + "'");
    if (it->second >= 0 && (int)a.size() != it->second) fail("wrong argument count for '" + f + "'");
    if (it->second < 0 && a.empty()) fail("'" + f + "' needs arguments");
    if (f == "if") return mk(Op::Tern, std::move(a));
    auto n = mk(Op::Call, std::move(a));
    n->fname = (f == "ln") ? "log" : (f == "sgn") ? "sign" : f;
    return n;
  }
};

inline double eval(const Node* n, const double* v) {
  switch (n->op) {
    case Op::Num: return n->num;
    case Op::Var: return v[n->var];
    case Op::Neg: return -eval(n->kids[0].get(), v);
    case Op::Not: return eval(n->kids[0].get(), v) == 0.0 ? 1.0 : 0.0;
    case Op::Add: return eval(n->kids[0].get(), v) + eval(n->kids[1].get(), v);
    case Op::Sub: return eval(n->kids[0].get(), v) - eval(n->kids[1].get(), v);
    case Op::Mul: return eval(n->kids[0].get(), v) * eval(n->kids[1].get(), v);
    case Op::Div: return eval(n->kids[0].get(), v) / eval(n->kids[1].get(), v);
    case Op::Mod: return std::fmod(eval(n->kids[0].get(), v), eval(n->kids[1].get(), v));
    case Op::Pow: return std::pow(eval(n->kids[0].get(), v), eval(n->kids[1].get(), v));
    case Op::Lt: return eval(n->kids[0].get(), v) < eval(n->kids[1].get(), v) ? 1.0 : 0.0;
    case Op::Le: return eval(n->kids[0].get(), v) <= eval(n->kids[1].get(), v) ? 1.0 : 0.0;
    case Op::Gt: return eval(n->kids[0].get(), v) > eval(n->kids[1].get(), v) ? 1.0 : 0.0;
    case Op::Ge: return eval(n->kids[0].get(), v) >= eval(n->kids[1].get(), v) ? 1.0 : 0.0;
    case Op::Eq: return eval(n->kids[0].get(), v) == eval(n->kids[1].get(), v) ? 1.0 : 0.0;
    case Op::Ne: return eval(n->kids[0].get(), v) != eval(n->kids[1].get(), v) ? 1.0 : 0.0;
    case Op::And: return (eval(n->kids[0].get(), v) != 0.0 && eval(n->kids[1].get(), v) != 0.0) ? 1.0 : 0.0;
    case Op::Or: return (eval(n->kids[0].get(), v) != 0.0 || eval(n->kids[1].get(), v) != 0.0) ? 1.0 : 0.0;
    case Op::Tern: return eval(n->kids[0].get(), v) != 0.0 ? eval(n->kids[1].get(), v) : eval(n->kids[2].get(), v);
    case Op::Table: {
      // first point not below the argument; the end values outside
      const Table& T = *n->table;
      const double a = eval(n->kids[0].get(), v);
      const size_t i = std::lower_bound(T.x.begin(), T.x.end(), a) - T.x.begin();
      if (i == 0) return T.y.front();
      if (i == T.x.size()) return T.y.back();
      const double w = (a - T.x[i - 1]) / (T.x[i] - T.x[i - 1]);
      return T.y[i - 1] + w * (T.y[i] - T.y[i - 1]);
    }
    case Op::Call: {
      const std::string& f = n->fname;
      double a = eval(n->kids[0].get(), v);
      if (f == "min" || f == "max") {
        for (size_t i = 1; i < n->kids.size(); ++i) {
          double b = eval(n->kids[i].get(), v);
          a = (f == "min") ? std::fmin(a, b) : std::fmax(a, b);
        }
        return a;
      }
      if (f == "pow") return std::pow(a, eval(n->kids[1].get(), v));
      if (f == "atan2") return std::atan2(a, eval(n->kids[1].get(), v));
      if (f == "sin") return std::sin(a);
      if (f == "cos") return std::cos(a);
      if (f == "tan") return std::tan(a);
      if (f == "asin") return std::asin(a);
      if (f == "acos") return std::acos(a);
      if (f == "atan") return std::atan(a);
      if (f == "sinh") return std::sinh(a);
      if (f == "cosh") return std::cosh(a);
      if (f == "tanh") return std::tanh(a);
      if (f == "exp") return std::exp(a);
      if (f == "log") return std::log(a);
      if (f == "log10") return std::log10(a);
      if (f == "log2") return std::log2(a);
      if (f == "sqrt") return std::sqrt(a);
      if (f == "abs") return std::fabs(a);
      if (f == "floor") return std::floor(a);
      if (f == "ceil") return std::ceil(a);
      if (f == "round") return std::round(a);
      if (f == "sign") return (a > 0) - (a < 0);
      throw std::runtime_error("expr: bad function " + f);
    }
  }
  return 0;
}

// A compiled scalar expression.  `null` mirrors the reference's rule that an
// empty or all-zero expression means "term absent"
// (functor_factory_parser.impl.hh:119-122).
struct Expr {
  NodeP root;
  bool null() const { return !root; }
  double operator()(const double* v) const { return eval(root.get(), v); }
};

inline bool is_zero_literal(const std::string& e) {
  // matches the reference's zero regex "-?([0]+)?([\.]?)([0]+)?"
  std::string t = trim(e);
  if (t.empty()) return true;
  size_t i = 0;
  if (t[i] == '-') ++i;
  bool any = false;
  for (; i < t.size(); ++i) {
    if (t[i] == '0') { any = true; continue; }
    if (t[i] == '.') continue;
    return false;
  }
  return any || t == "." || t == "-";
}

inline Expr compile(const std::string& text, const Scope& scope) {
  Expr e;
  if (is_zero_literal(text)) return e;
  e.root = Parser(text, scope).parse();
  return e;
}

}  // namespace orc
