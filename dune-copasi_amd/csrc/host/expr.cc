#include "expr.hh"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <functional>
#include <sstream>

#include "error.hh"

namespace dcopasi::expr {

// ------------------------------------------------------------ construction
Ex constant(double v) {
  auto n = std::make_shared<Node>();
  n->kind = Kind::Const;
  n->value = v;
  return n;
}

Ex variable(const std::string& name) {
  auto n = std::make_shared<Node>();
  n->kind = Kind::Var;
  n->name = name;
  return n;
}

bool is_const(const Ex& e, double* v) {
  if (!e) {
    if (v) *v = 0.0;
    return true;
  }
  if (e->kind != Kind::Const) return false;
  if (v) *v = e->value;
  return true;
}

static double apply_fn(const std::string& f, const std::vector<double>& a) {
  if (f == "min" || f == "max") {
    double r = a[0];
    for (size_t i = 1; i < a.size(); ++i) r = (f == "min") ? std::fmin(r, a[i]) : std::fmax(r, a[i]);
    return r;
  }
  if (f == "pow") return std::pow(a[0], a[1]);
  if (f == "atan2") return std::atan2(a[0], a[1]);
  const double x = a[0];
  if (f == "sin") return std::sin(x);
  if (f == "cos") return std::cos(x);
  if (f == "tan") return std::tan(x);
  if (f == "asin") return std::asin(x);
  if (f == "acos") return std::acos(x);
  if (f == "atan") return std::atan(x);
  if (f == "sinh") return std::sinh(x);
  if (f == "cosh") return std::cosh(x);
  if (f == "tanh") return std::tanh(x);
  if (f == "exp") return std::exp(x);
  if (f == "log") return std::log(x);
  if (f == "log10") return std::log10(x);
  if (f == "log2") return std::log2(x);
  if (f == "sqrt") return std::sqrt(x);
  if (f == "abs") return std::fabs(x);
  if (f == "floor") return std::floor(x);
  if (f == "ceil") return std::ceil(x);
  if (f == "round") return std::round(x);
  if (f == "trunc") return std::trunc(x);
  if (f == "sign") return double((x > 0) - (x < 0));
  throw IOError("unknown function '" + f + "'");
}

static double fold(Kind k, const std::vector<double>& a) {
  switch (k) {
    case Kind::Neg: return -a[0];
    case Kind::Not: return a[0] == 0.0 ? 1.0 : 0.0;
    case Kind::Add: return a[0] + a[1];
    case Kind::Sub: return a[0] - a[1];
    case Kind::Mul: return a[0] * a[1];
    case Kind::Div: return a[0] / a[1];
    case Kind::Pow: return std::pow(a[0], a[1]);
    case Kind::Mod: return std::fmod(a[0], a[1]);
    case Kind::Lt: return a[0] < a[1];
    case Kind::Le: return a[0] <= a[1];
    case Kind::Gt: return a[0] > a[1];
    case Kind::Ge: return a[0] >= a[1];
    case Kind::Eq: return a[0] == a[1];
    case Kind::Ne: return a[0] != a[1];
    case Kind::And: return (a[0] != 0.0 && a[1] != 0.0) ? 1.0 : 0.0;
    case Kind::Or: return (a[0] != 0.0 || a[1] != 0.0) ? 1.0 : 0.0;
    case Kind::Cond: return a[0] != 0.0 ? a[1] : a[2];
    default: break;
  }
  throw Error("expr: cannot fold");
}

Ex make(Kind k, std::vector<Ex> args) {
  for (auto& a : args)
    if (!a) a = constant(0.0);
  bool all_const = true;
  std::vector<double> cv;
  for (auto& a : args) {
    double v;
    if (is_const(a, &v)) cv.push_back(v);
    else all_const = false;
  }
  if (all_const) return constant(fold(k, cv));
  double v;
  switch (k) {
    case Kind::Add:
      if (is_const(args[0], &v) && v == 0.0) return args[1];
      if (is_const(args[1], &v) && v == 0.0) return args[0];
      break;
    case Kind::Sub:
      if (is_const(args[1], &v) && v == 0.0) return args[0];
      if (is_const(args[0], &v) && v == 0.0) return make(Kind::Neg, {args[1]});
      break;
    case Kind::Mul:
      if ((is_const(args[0], &v) && v == 0.0) || (is_const(args[1], &v) && v == 0.0)) return constant(0.0);
      if (is_const(args[0], &v) && v == 1.0) return args[1];
      if (is_const(args[1], &v) && v == 1.0) return args[0];
      break;
    case Kind::Div:
      if (is_const(args[0], &v) && v == 0.0) return constant(0.0);
      if (is_const(args[1], &v) && v == 1.0) return args[0];
      break;
    case Kind::Pow:
      if (is_const(args[1], &v) && v == 1.0) return args[0];
      if (is_const(args[1], &v) && v == 0.0) return constant(1.0);
      break;
    case Kind::Neg:
      if (args[0]->kind == Kind::Neg) return args[0]->args[0];
      break;
    case Kind::Cond:
      if (is_const(args[0], &v)) return v != 0.0 ? args[1] : args[2];
      break;
    default: break;
  }
  auto n = std::make_shared<Node>();
  n->kind = k;
  n->args = std::move(args);
  return n;
}

Ex call(const std::string& fn, std::vector<Ex> args) {
  for (auto& a : args)
    if (!a) a = constant(0.0);
  bool all_const = true;
  std::vector<double> cv;
  for (auto& a : args) {
    double v;
    if (is_const(a, &v)) cv.push_back(v);
    else all_const = false;
  }
  if (all_const) return constant(apply_fn(fn, cv));
  auto n = std::make_shared<Node>();
  n->kind = Kind::Call;
  n->name = fn;
  n->args = std::move(args);
  return n;
}

// ------------------------------------------------------------ parsing
namespace {

struct Token {
  enum Type { End, Number, Name, Sym } type;
  std::string text;
  double number = 0;
};

std::vector<Token> tokenize(const std::string& s) {
  std::vector<Token> out;
  size_t i = 0;
  while (i < s.size()) {
    char c = s[i];
    if (std::isspace(static_cast<unsigned char>(c))) {
      ++i;
      continue;
    }
    if (std::isdigit(static_cast<unsigned char>(c)) ||
        (c == '.' && i + 1 < s.size() && std::isdigit(static_cast<unsigned char>(s[i + 1])))) {
      const char* begin = s.c_str() + i;
      char* end = nullptr;
      double v = std::strtod(begin, &end);
      out.push_back({Token::Number, std::string(begin, static_cast<size_t>(end - begin)), v});
      i += static_cast<size_t>(end - begin);
      continue;
    }
    if (std::isalpha(static_cast<unsigned char>(c)) || c == '_') {
      size_t j = i;
      while (j < s.size() && (std::isalnum(static_cast<unsigned char>(s[j])) || s[j] == '_')) ++j;
      std::string w = s.substr(i, j - i);
      if (w == "and" || w == "or" || w == "not") out.push_back({Token::Sym, w == "and" ? "&&" : w == "or" ? "||" : "!"});
      else out.push_back({Token::Name, w});
      i = j;
      continue;
    }
    static const char* two[] = {"<=", ">=", "==", "!=", "&&", "||"};
    bool matched = false;
    for (const char* t : two)
      if (s.compare(i, 2, t) == 0) {
        out.push_back({Token::Sym, t});
        i += 2;
        matched = true;
        break;
      }
    if (matched) continue;
    if (std::string("+-*/^%<>!?:(),").find(c) == std::string::npos)
      throw IOError(std::string("expression '") + s + "': unexpected character '" + c + "'");
    out.push_back({Token::Sym, std::string(1, c)});
    ++i;
  }
  out.push_back({Token::End, ""});
  return out;
}

class Reader {
 public:
  Reader(const std::string& src, const Context& ctx, const std::map<std::string, Ex>& locals, int depth)
      : src_(src), toks_(tokenize(src)), ctx_(ctx), locals_(locals), depth_(depth) {}

  Ex expression() {
    Ex e = conditional();
    if (peek().type != Token::End) error("unexpected '" + peek().text + "'");
    return e;
  }

 private:
  const std::string src_;
  std::vector<Token> toks_;
  size_t at_ = 0;
  const Context& ctx_;
  const std::map<std::string, Ex>& locals_;
  int depth_;

  [[noreturn]] void error(const std::string& what) const { throw IOError("expression '" + src_ + "': " + what); }
  const Token& peek() const { return toks_[at_]; }
  bool accept(const char* sym) {
    if (peek().type == Token::Sym && peek().text == sym) {
      ++at_;
      return true;
    }
    return false;
  }

  // binary levels, lowest first
  struct Level {
    std::vector<std::pair<const char*, Kind>> ops;
  };
  static const std::vector<Level>& levels() {
    static const std::vector<Level> L = {
        {{{"||", Kind::Or}}},
        {{{"&&", Kind::And}}},
        {{{"==", Kind::Eq}, {"!=", Kind::Ne}}},
        {{{"<=", Kind::Le}, {">=", Kind::Ge}, {"<", Kind::Lt}, {">", Kind::Gt}}},
        {{{"+", Kind::Add}, {"-", Kind::Sub}}},
        {{{"*", Kind::Mul}, {"/", Kind::Div}, {"%", Kind::Mod}}},
    };
    return L;
  }

  Ex conditional() {
    Ex c = binary(0);
    if (accept("?")) {
      Ex a = conditional();
      if (!accept(":")) error("expected ':' in conditional");
      Ex b = conditional();
      return make(Kind::Cond, {c, a, b});
    }
    return c;
  }

  Ex binary(size_t level) {
    if (level == levels().size()) return prefix();
    Ex lhs = binary(level + 1);
    for (;;) {
      bool found = false;
      for (auto& [sym, kind] : levels()[level].ops) {
        if (accept(sym)) {
          lhs = make(kind, {lhs, binary(level + 1)});
          found = true;
          break;
        }
      }
      if (!found) return lhs;
    }
  }

  Ex prefix() {
    if (accept("-")) return make(Kind::Neg, {prefix()});
    if (accept("+")) return prefix();
    if (accept("!")) return make(Kind::Not, {prefix()});
    Ex base = atom();
    if (accept("^")) return make(Kind::Pow, {base, prefix()});
    return base;
  }

  Ex atom() {
    const Token t = peek();
    if (t.type == Token::Number) {
      ++at_;
      return constant(t.number);
    }
    if (accept("(")) {
      Ex e = conditional();
      if (!accept(")")) error("expected ')'");
      return e;
    }
    if (t.type != Token::Name) error("expected a value, found '" + t.text + "'");
    ++at_;
    if (accept("(")) {
      std::vector<Ex> args;
      if (!accept(")")) {
        do args.push_back(conditional());
        while (accept(","));
        if (!accept(")")) error("expected ')' after arguments of '" + t.text + "'");
      }
      return apply(t.text, std::move(args));
    }
    if (auto it = locals_.find(t.text); it != locals_.end()) return it->second;
    if (t.text == "no_value") return constant(DBL_MAX);
    if (auto it = ctx_.constants.find(t.text); it != ctx_.constants.end())
      return ctx_.fold_constants ? constant(it->second) : variable("$" + t.text);
    return variable(t.text);
  }

  Ex apply(const std::string& f, std::vector<Ex> args) {
    if (auto it = ctx_.functions.find(f); it != ctx_.functions.end()) {
      const Function& fn = it->second;
      if (fn.params.size() != args.size()) error("function '" + f + "' takes " + std::to_string(fn.params.size()) + " arguments");
      if (depth_ > 32) error("function expansion too deep");
      std::map<std::string, Ex> bound;
      for (size_t i = 0; i < args.size(); ++i) bound[fn.params[i]] = args[i];
      return Reader(fn.body, ctx_, bound, depth_ + 1).expression();
    }
    std::string name = f == "ln" ? "log" : f == "sgn" ? "sign" : f;
    if (name == "if") {
      if (args.size() != 3) error("if() takes 3 arguments");
      return make(Kind::Cond, std::move(args));
    }
    static const std::map<std::string, int> arity = {
        {"sin", 1}, {"cos", 1}, {"tan", 1}, {"asin", 1}, {"acos", 1}, {"atan", 1}, {"sinh", 1},
        {"cosh", 1}, {"tanh", 1}, {"exp", 1}, {"log", 1}, {"log10", 1}, {"log2", 1}, {"sqrt", 1},
        {"abs", 1}, {"floor", 1}, {"ceil", 1}, {"round", 1}, {"sign", 1}, {"pow", 2}, {"atan2", 2},
        {"min", -1}, {"max", -1}};
    auto it = arity.find(name);
    if (it == arity.end()) error("unknown function '" + f + "'");
    if (it->second >= 0 && static_cast<int>(args.size()) != it->second) error("wrong number of arguments to '" + f + "'");
    if (args.empty()) error("'" + f + "' needs arguments");
    if (it->second < 0 && args.size() > 2) {  // fold n-ary min/max into a binary chain
      Ex acc = args[0];
      for (size_t i = 1; i < args.size(); ++i) acc = call(name, {acc, args[i]});
      return acc;
    }
    if (it->second < 0 && args.size() == 1) return args[0];
    return call(name, std::move(args));
  }
};

bool zero_literal(const std::string& text) {
  // the reference's "-?([0]+)?([\.]?)([0]+)?" test
  size_t i = 0, n = text.size();
  while (i < n && std::isspace(static_cast<unsigned char>(text[i]))) ++i;
  while (n > i && std::isspace(static_cast<unsigned char>(text[n - 1]))) --n;
  if (i < n && text[i] == '-') ++i;
  int dots = 0;
  for (; i < n; ++i) {
    if (text[i] == '0') continue;
    if (text[i] == '.' && ++dots == 1) continue;
    return false;
  }
  return true;
}

}  // namespace

Ex parse(const std::string& text, const Context& ctx) {
  if (zero_literal(text)) return nullptr;
  static const std::map<std::string, Ex> none;
  Ex e = Reader(text, ctx, none, 0).expression();
  return e;
}

// ------------------------------------------------------------ analysis
bool depends_on(const Ex& e, const std::string& var) {
  if (!e) return false;
  if (e->kind == Kind::Var) return e->name == var;
  for (auto& a : e->args)
    if (depends_on(a, var)) return true;
  return false;
}

void collect_vars(const Ex& e, std::vector<std::string>& out) {
  if (!e) return;
  if (e->kind == Kind::Var) {
    for (auto& s : out)
      if (s == e->name) return;
    out.push_back(e->name);
    return;
  }
  for (auto& a : e->args) collect_vars(a, out);
}

// ------------------------------------------------------------ differentiation
Ex derivative(const Ex& e, const std::string& x) {
  if (!e || !depends_on(e, x)) return nullptr;
  auto D = [&](const Ex& a) { return derivative(a, x); };
  auto C = [](double v) { return constant(v); };
  const auto& A = e->args;
  Ex r;
  switch (e->kind) {
    case Kind::Const: return nullptr;
    case Kind::Var: return C(1.0);
    case Kind::Neg: r = make(Kind::Neg, {D(A[0])}); break;
    case Kind::Add: r = make(Kind::Add, {D(A[0]), D(A[1])}); break;
    case Kind::Sub: r = make(Kind::Sub, {D(A[0]), D(A[1])}); break;
    case Kind::Mul:
      r = make(Kind::Add, {make(Kind::Mul, {D(A[0]), A[1]}), make(Kind::Mul, {A[0], D(A[1])})});
      break;
    case Kind::Div:
      r = make(Kind::Div, {make(Kind::Sub, {make(Kind::Mul, {D(A[0]), A[1]}), make(Kind::Mul, {A[0], D(A[1])})}),
                           make(Kind::Mul, {A[1], A[1]})});
      break;
    case Kind::Pow: {
      double p;
      if (is_const(A[1], &p)) {
        r = make(Kind::Mul, {make(Kind::Mul, {C(p), make(Kind::Pow, {A[0], C(p - 1.0)})}), D(A[0])});
      } else {
        // a^b (b' ln a + b a'/a)
        Ex t1 = make(Kind::Mul, {D(A[1]), call("log", {A[0]})});
        Ex t2 = make(Kind::Div, {make(Kind::Mul, {A[1], D(A[0])}), A[0]});
        r = make(Kind::Mul, {e, make(Kind::Add, {t1, t2})});
      }
      break;
    }
    case Kind::Mod:
      // fmod(a, b) = a - trunc(a / b) b: the quotient is piecewise constant
      if (depends_on(A[1], x))
        r = make(Kind::Sub, {D(A[0]), make(Kind::Mul, {call("trunc", {make(Kind::Div, {A[0], A[1]})}), D(A[1])})});
      else
        r = D(A[0]);
      break;
    case Kind::Cond: r = make(Kind::Cond, {A[0], D(A[1]), D(A[2])}); break;
    case Kind::Lt: case Kind::Le: case Kind::Gt: case Kind::Ge: case Kind::Eq: case Kind::Ne:
    case Kind::And: case Kind::Or: case Kind::Not:
      return nullptr;  // piecewise constant
    case Kind::Call: {
      const std::string& f = e->name;
      const Ex& a = A[0];
      Ex da = D(a);
      if (f == "sin") r = make(Kind::Mul, {call("cos", {a}), da});
      else if (f == "cos") r = make(Kind::Neg, {make(Kind::Mul, {call("sin", {a}), da})});
      else if (f == "tan") r = make(Kind::Mul, {make(Kind::Add, {C(1.0), make(Kind::Mul, {e, e})}), da});
      else if (f == "asin") r = make(Kind::Div, {da, call("sqrt", {make(Kind::Sub, {C(1.0), make(Kind::Mul, {a, a})})})});
      else if (f == "acos") r = make(Kind::Neg, {make(Kind::Div, {da, call("sqrt", {make(Kind::Sub, {C(1.0), make(Kind::Mul, {a, a})})})})});
      else if (f == "atan") r = make(Kind::Div, {da, make(Kind::Add, {C(1.0), make(Kind::Mul, {a, a})})});
      else if (f == "sinh") r = make(Kind::Mul, {call("cosh", {a}), da});
      else if (f == "cosh") r = make(Kind::Mul, {call("sinh", {a}), da});
      else if (f == "tanh") r = make(Kind::Mul, {make(Kind::Sub, {C(1.0), make(Kind::Mul, {e, e})}), da});
      else if (f == "exp") r = make(Kind::Mul, {e, da});
      else if (f == "log") r = make(Kind::Div, {da, a});
      else if (f == "log10") r = make(Kind::Div, {da, make(Kind::Mul, {a, C(std::log(10.0))})});
      else if (f == "log2") r = make(Kind::Div, {da, make(Kind::Mul, {a, C(std::log(2.0))})});
      else if (f == "sqrt") r = make(Kind::Div, {da, make(Kind::Mul, {C(2.0), e})});
      else if (f == "abs") r = make(Kind::Mul, {call("sign", {a}), da});
      else if (f == "floor" || f == "ceil" || f == "round" || f == "sign" || f == "trunc") return nullptr;
      else if (f == "pow") r = derivative(make(Kind::Pow, {A[0], A[1]}), x);
      else if (f == "atan2") {
        // d atan2(y, x) = (x dy - y dx) / (x^2 + y^2)
        const Ex& y = A[0];
        const Ex& xx = A[1];
        r = make(Kind::Div, {make(Kind::Sub, {make(Kind::Mul, {xx, D(y)}), make(Kind::Mul, {y, D(xx)})}),
                             make(Kind::Add, {make(Kind::Mul, {xx, xx}), make(Kind::Mul, {y, y})})});
      } else if (f == "min") r = make(Kind::Cond, {make(Kind::Le, {A[0], A[1]}), D(A[0]), D(A[1])});
      else if (f == "max") r = make(Kind::Cond, {make(Kind::Ge, {A[0], A[1]}), D(A[0]), D(A[1])});
      else throw NotImplemented("no derivative rule for '" + f + "'");
      break;
    }
  }
  double v;
  if (is_const(r, &v) && v == 0.0) return nullptr;
  return r;
}

// ------------------------------------------------------------ emission
static std::string num(double v) {
  if (v == DBL_MAX) return "1.7976931348623157e308";
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.17g", v);
  std::string s = buf;
  if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
  return "(" + s + ")";
}

std::string emit(const Ex& e, const std::map<std::string, std::string>& map) {
  if (!e) return "0.0";
  auto E = [&](int i) { return emit(e->args[i], map); };
  switch (e->kind) {
    case Kind::Const: return num(e->value);
    case Kind::Var: {
      auto it = map.find(e->name);
      if (it == map.end()) throw IOError("unknown symbol '" + e->name + "' in expression");
      return it->second;
    }
    case Kind::Neg: return "(-" + E(0) + ")";
    case Kind::Not: return "((" + E(0) + ") == 0.0 ? 1.0 : 0.0)";
    case Kind::Add: return "(" + E(0) + " + " + E(1) + ")";
    case Kind::Sub: return "(" + E(0) + " - " + E(1) + ")";
    case Kind::Mul: return "(" + E(0) + " * " + E(1) + ")";
    case Kind::Div: return "(" + E(0) + " / " + E(1) + ")";
    case Kind::Mod: return "fmod(" + E(0) + ", " + E(1) + ")";
    case Kind::Pow: {
      double p;
      if (is_const(e->args[1], &p) && (p == 2.0 || p == 3.0)) {
        std::string b = E(0);
        return p == 2.0 ? "(" + b + " * " + b + ")" : "(" + b + " * " + b + " * " + b + ")";
      }
      return "pow(" + E(0) + ", " + E(1) + ")";
    }
    case Kind::Lt: return "((" + E(0) + " < " + E(1) + ") ? 1.0 : 0.0)";
    case Kind::Le: return "((" + E(0) + " <= " + E(1) + ") ? 1.0 : 0.0)";
    case Kind::Gt: return "((" + E(0) + " > " + E(1) + ") ? 1.0 : 0.0)";
    case Kind::Ge: return "((" + E(0) + " >= " + E(1) + ") ? 1.0 : 0.0)";
    case Kind::Eq: return "((" + E(0) + " == " + E(1) + ") ? 1.0 : 0.0)";
    case Kind::Ne: return "((" + E(0) + " != " + E(1) + ") ? 1.0 : 0.0)";
    case Kind::And: return "(((" + E(0) + ") != 0.0 && (" + E(1) + ") != 0.0) ? 1.0 : 0.0)";
    case Kind::Or: return "(((" + E(0) + ") != 0.0 || (" + E(1) + ") != 0.0) ? 1.0 : 0.0)";
    case Kind::Cond: return "((" + E(0) + ") != 0.0 ? " + E(1) + " : " + E(2) + ")";
    case Kind::Call: {
      const std::string& f = e->name;
      if (f == "min") return "fmin(" + E(0) + ", " + E(1) + ")";
      if (f == "max") return "fmax(" + E(0) + ", " + E(1) + ")";
      if (f == "abs") return "fabs(" + E(0) + ")";
      if (f == "sign") return "dcg_sign(" + E(0) + ")";
      std::string s = f + "(";
      for (size_t i = 0; i < e->args.size(); ++i) s += (i ? ", " : "") + E(static_cast<int>(i));
      return s + ")";
    }
  }
  return "0.0";
}

// ------------------------------------------------------------ host evaluation
double evaluate(const Ex& e, const std::map<std::string, double>& vars) {
  if (!e) return 0.0;
  if (e->kind == Kind::Const) return e->value;
  if (e->kind == Kind::Var) {
    auto it = vars.find(e->name);
    if (it == vars.end()) throw IOError("unknown symbol '" + e->name + "' in expression");
    return it->second;
  }
  if (e->kind == Kind::Cond) return evaluate(e->args[0], vars) != 0.0 ? evaluate(e->args[1], vars) : evaluate(e->args[2], vars);
  std::vector<double> a;
  for (auto& x : e->args) a.push_back(evaluate(x, vars));
  if (e->kind == Kind::Call) return apply_fn(e->name, a);
  return fold(e->kind, a);
}

// postfix program: opcode stream; operands: slot or constant index
namespace {
enum Code : int { PUSH_CONST = 1000, PUSH_SLOT, OP_KIND, OP_CALL };
const std::vector<std::string>& fn_names() {
  static const std::vector<std::string> n = {"sin", "cos", "tan", "asin", "acos", "atan", "sinh", "cosh",
                                             "tanh", "exp", "log", "log10", "log2", "sqrt", "abs", "floor",
                                             "ceil", "round", "sign", "pow", "atan2", "min", "max", "trunc"};
  return n;
}
void compile_node(const Ex& e, const std::map<std::string, int>& slots, Bound& b) {
  if (!e) {
    b.code.push_back(PUSH_CONST);
    b.code.push_back(static_cast<int>(b.consts.size()));
    b.consts.push_back(0.0);
    return;
  }
  switch (e->kind) {
    case Kind::Const:
      b.code.push_back(PUSH_CONST);
      b.code.push_back(static_cast<int>(b.consts.size()));
      b.consts.push_back(e->value);
      return;
    case Kind::Var: {
      auto it = slots.find(e->name);
      if (it == slots.end()) throw IOError("unknown symbol '" + e->name + "' in expression");
      b.code.push_back(PUSH_SLOT);
      b.code.push_back(it->second);
      return;
    }
    default: break;
  }
  for (auto& a : e->args) compile_node(a, slots, b);
  if (e->kind == Kind::Call) {
    const auto& n = fn_names();
    int id = -1;
    for (size_t i = 0; i < n.size(); ++i)
      if (n[i] == e->name) id = static_cast<int>(i);
    b.code.push_back(OP_CALL);
    b.code.push_back(id);
    b.code.push_back(static_cast<int>(e->args.size()));
  } else {
    b.code.push_back(OP_KIND);
    b.code.push_back(static_cast<int>(e->kind));
    b.code.push_back(static_cast<int>(e->args.size()));
  }
}
}  // namespace

Bound bind(const Ex& e, const std::map<std::string, int>& slots) {
  Bound b;
  compile_node(e, slots, b);
  return b;
}

double Bound::operator()(const double* s) const {
  double stack[64];
  int sp = 0;
  std::vector<double> args;
  for (size_t pc = 0; pc < code.size();) {
    int op = code[pc++];
    if (op == PUSH_CONST) {
      stack[sp++] = consts[code[pc++]];
    } else if (op == PUSH_SLOT) {
      stack[sp++] = s[code[pc++]];
    } else {
      int what = code[pc++];
      int n = code[pc++];
      args.assign(stack + sp - n, stack + sp);
      sp -= n;
      stack[sp++] = (op == OP_CALL) ? apply_fn(fn_names()[what], args) : fold(static_cast<Kind>(what), args);
    }
    if (sp >= 63) throw Error("expression too deep for host evaluation");
  }
  return stack[0];
}

}  // namespace dcopasi::expr
