#include "spec.hh"

#include <algorithm>
#include <cstdio>

#include "error.hh"

namespace dcopasi {

static const char* kAxis = "xyz";

static expr::Context read_context(const ParameterTree& cfg) {
  expr::Context ctx;
  // parser_context.<name>.type = constant | function | interpolation (context.cc:56-173)
  for (const auto& name : cfg.sub_keys("parser_context")) {
    const std::string p = "parser_context." + name;
    const std::string type = cfg.get(p + ".type", std::string());
    if (type == "constant") {
      ctx.constants[name] = cfg.get(p + ".value", 0.0);
    } else if (type == "function") {
      const std::string& text = cfg.at(p + ".expression");
      auto colon = text.find(':');
      if (colon == std::string::npos) throw IOError("parser_context." + name + ": expected 'args: expression'");
      expr::Function f;
      std::string args = text.substr(0, colon);
      size_t start = 0;
      for (;;) {
        size_t comma = args.find(',', start);
        std::string a = args.substr(start, comma == std::string::npos ? std::string::npos : comma - start);
        size_t b = a.find_first_not_of(" \t"), e = a.find_last_not_of(" \t");
        if (b != std::string::npos) f.params.push_back(a.substr(b, e - b + 1));
        if (comma == std::string::npos) break;
        start = comma + 1;
      }
      f.body = text.substr(colon + 1);
      ctx.functions[name] = f;
    } else if (type == "interpolation") {
      // context.cc:72-97: linear between the (domain, range) points, the end
      // values outside; inlined as a one-argument function of nested ifs
      const std::vector<double> dom = cfg.get_list(p + ".domain"), rng = cfg.get_list(p + ".range");
      if (!std::is_sorted(dom.begin(), dom.end())) throw IOError("The interpolation domain must be sorted");
      if (dom.size() < 2 || dom.size() != rng.size())
        throw IOError("Interpolation range and domain must have at least two points and be the same size");
      auto num = [](double v) {
        char b[40];
        std::snprintf(b, sizeof b, "(%.17g)", v);
        return std::string(b);
      };
      std::string body = num(rng.back());
      for (size_t i = dom.size(); i-- > 1;) {
        if (dom[i] == dom[i - 1]) continue;  // an empty interval
        body = "if(x <= " + num(dom[i]) + ", " + num(rng[i - 1]) + " + ((x - " + num(dom[i - 1]) + ") / (" +
               num(dom[i]) + " - " + num(dom[i - 1]) + ")) * (" + num(rng[i]) + " - " + num(rng[i - 1]) + "), " +
               body + ")";
      }
      body = "if(x <= " + num(dom[0]) + ", " + num(rng[0]) + ", " + body + ")";
      expr::Function f;
      f.params = {"x"};
      f.body = body;
      ctx.functions[name] = f;
    } else if (!type.empty()) {
      throw NotImplemented("parser_context type '" + type + "' (" + name + ")");
    }
  }
  return ctx;
}

static Ex expression(const ParameterTree& cfg, const std::string& key, const expr::Context& ctx) {
  return expr::parse(cfg.get(key, std::string()), ctx);
}

static bool read_tensor(const ParameterTree& cfg, const std::string& p, int dim, const expr::Context& ctx,
                        TensorTerm& t) {
  const std::string type = cfg.get(p + ".type", std::string("scalar"));
  bool any = false;
  if (type == "scalar") {
    t.tensor = false;
    t.d[0][0] = expression(cfg, p + ".expression", ctx);
    any = t.d[0][0] != nullptr;
  } else if (type == "tensor") {
    t.tensor = true;
    for (int i = 0; i < dim; ++i)
      for (int j = 0; j < dim; ++j) {
        t.d[i][j] = expression(cfg, p + "." + kAxis[i] + kAxis[j] + ".expression", ctx);
        any = any || t.d[i][j] != nullptr;
      }
  } else {
    throw IOError("not known type '" + p + ".type = " + type + "'");
  }
  return any;
}

static std::vector<Partial> partials(const ParameterTree& cfg, const std::string& prefix, const ModelSpec& m,
                                     const std::vector<int>& candidates) {
  std::vector<Partial> out;
  for (int s : candidates) {
    Ex e = expression(cfg, prefix + ".jacobian." + m.species[s].name + ".expression", m.ctx);
    if (e) out.push_back({s, e});
  }
  return out;
}

// d(term)/d(species) for every species in `candidates` the term depends on
static std::vector<Partial> symbolic(const Ex& term, const ModelSpec& m, const std::vector<int>& candidates) {
  std::vector<Partial> out;
  for (int s : candidates)
    if (Ex d = expr::derivative(term, m.species[s].name)) out.push_back({s, d});
  return out;
}

ModelSpec read_model(const ParameterTree& cfg, int dim, bool fold_constants) {
  ModelSpec m;
  m.dim = dim;
  m.ctx = read_context(cfg);
  // compartment and initial expressions are evaluated on the host: always folded
  const expr::Context folded = m.ctx;
  m.ctx.fold_constants = fold_constants;
  if (!fold_constants)
    for (const auto& [name, value] : m.ctx.constants) {
      m.params.push_back(name);
      m.param_values.push_back(value);
    }
  // compartments in INI order (make_multi_domain_grid.hh:128-160)
  for (const auto& c : cfg.sub_keys("compartments")) {
    const std::string type = cfg.get("compartments." + c + ".type", std::string("expression"));
    if (type != "expression") throw NotImplemented("Not know type '" + type + "' of compartment");
    m.comp_names.push_back(c);
    m.comp_expr.push_back(expression(cfg, "compartments." + c + ".expression", folded));
  }
  m.comp_species.resize(m.comp_names.size());
  for (const auto& s : cfg.sub_keys("model.scalar_field")) {
    const std::string cname = cfg.get("model.scalar_field." + s + ".compartment", std::string());
    if (cname.empty()) continue;
    int cid = -1;
    for (size_t i = 0; i < m.comp_names.size(); ++i)
      if (m.comp_names[i] == cname) cid = static_cast<int>(i);
    if (cid < 0) throw IOError("scalar field '" + s + "' lives in unknown compartment '" + cname + "'");
    for (const auto& other : m.species)
      if (other.name == s) throw IOError("Variable with name '" + s + "' is repeated");
    SpeciesSpec sp;
    sp.name = s;
    sp.comp = cid;
    sp.local = static_cast<int>(m.comp_species[cid].size());
    m.comp_species[cid].push_back(static_cast<int>(m.species.size()));
    m.species.push_back(sp);
  }
  if (m.species.empty())
    throw IOError("Basis has dimension 0, make sure to have at least one 'scalar_field' with a non-empty 'compartment'");

  const std::string jt = cfg.get("model.jacobian.type", std::string("analytical"));
  if (jt == "analytical") m.jacobian = JacobianType::Analytical;
  else if (jt == "numerical") m.jacobian = JacobianType::Numerical;
  else if (jt == "symbolic") m.jacobian = JacobianType::Symbolic;
  else throw IOError("The option 'model.jacobian.type' must be either 'analytical', 'numerical' or 'symbolic'");
  m.fd_epsilon = cfg.get("model.jacobian.epsilon", 1e-7);
  m.is_linear = cfg.get("model.is_linear", false);
  if (m.is_linear) m.jacobian = JacobianType::Analytical;  // local_operator.hh:252-253

  std::vector<int> all(m.species.size());
  for (size_t i = 0; i < all.size(); ++i) all[i] = static_cast<int>(i);
  const bool sym = m.jacobian == JacobianType::Symbolic;
  for (auto& sp : m.species) {
    const std::string p = "model.scalar_field." + sp.name;
    const auto& same = m.comp_species[sp.comp];
    sp.storage = expression(cfg, p + ".storage.expression", m.ctx);
    if (sp.storage) sp.storage_jac = sym ? symbolic(sp.storage, m, same) : partials(cfg, p + ".storage", m, same);
    sp.reaction = expression(cfg, p + ".reaction.expression", m.ctx);
    if (sp.reaction) sp.reaction_jac = sym ? symbolic(sp.reaction, m, same) : partials(cfg, p + ".reaction", m, same);
    bool has_v = false;
    for (int k = 0; k < dim; ++k) {
      sp.velocity[k] = expression(cfg, p + ".velocity." + kAxis[k] + ".expression", m.ctx);
      has_v = has_v || sp.velocity[k];
    }
    if (has_v) {
      for (int w : same) {
        std::array<Ex, 3> dv{};
        bool any = false;
        for (int k = 0; k < dim; ++k) {
          dv[k] = sym ? expr::derivative(sp.velocity[k], m.species[w].name)
                      : expression(cfg, p + ".velocity.jacobian." + m.species[w].name + "." + kAxis[k] + ".expression", m.ctx);
          any = any || dv[k];
        }
        if (any) sp.velocity_jac.push_back({w, dv});
      }
    }
    for (int w : same) {
      const std::string dp = p + ".cross_diffusion." + m.species[w].name;
      if (!cfg.has_sub(dp)) continue;
      DiffusionTerm term;
      term.wrt = w;
      if (!read_tensor(cfg, dp, dim, m.ctx, term.D)) continue;
      for (int v : same) {
        TensorTerm jt2;
        bool any = false;
        if (sym) {
          jt2.tensor = term.D.tensor;
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
              jt2.d[i][j] = expr::derivative(term.D.d[i][j], m.species[v].name);
              any = any || jt2.d[i][j];
            }
        } else if (cfg.has_sub(dp + ".jacobian." + m.species[v].name)) {
          any = read_tensor(cfg, dp + ".jacobian." + m.species[v].name, dim, m.ctx, jt2);
        }
        if (any) term.jac.push_back({v, jt2});
      }
      sp.diffusion.push_back(term);
    }
    for (size_t c = 0; c < m.comp_names.size(); ++c) {
      const std::string op = p + ".outflow." + m.comp_names[c];
      if (!cfg.has_sub(op)) continue;
      OutflowTerm o;
      o.to_comp = static_cast<int>(c);
      o.e = expression(cfg, op + ".expression", m.ctx);
      if (!o.e) continue;
      o.jac = sym ? symbolic(o.e, m, all) : partials(cfg, op, m, all);
      if (static_cast<int>(c) == sp.comp) m.has_outflow_boundary = true;
      else m.has_outflow_interface = true;
      sp.outflow.push_back(o);
    }
    m.initial.push_back(expression(cfg, p + ".initial.expression", folded));
  }
  return m;
}

}  // namespace dcopasi
