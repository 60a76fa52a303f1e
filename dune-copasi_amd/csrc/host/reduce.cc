#include "reduce.hh"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <sstream>

#include "codegen.hh"
#include "error.hh"
#include "model.hh"
#include "rtc.hh"

namespace dcopasi {

namespace {

std::string trim(const std::string& s) {
  size_t b = s.find_first_not_of(" \t\r\n"), e = s.find_last_not_of(" \t\r\n");
  return b == std::string::npos ? std::string() : s.substr(b, e - b + 1);
}

// "a, b: body" (parser/context.cc parse_function_expression)
ReduceFunction parse_function(const std::string& key, const std::string& text, size_t nargs,
                              const expr::Context& ctx) {
  const size_t colon = text.find(':');
  if (colon == std::string::npos) throw IOError(key + ": expected 'arguments: expression'");
  ReduceFunction f;
  std::stringstream ss(text.substr(0, colon));
  std::string a;
  while (std::getline(ss, a, ',')) {
    a = trim(a);
    if (!a.empty()) f.args.push_back(a);
  }
  if (f.args.size() != nargs) {
    if (nargs == 2) throw IOError("Reduction arguments must be exactly 2 (" + key + ")");
    throw IOError(key + ": function must have exactly 1 argument");
  }
  // the arguments shadow constants of the same name
  expr::Context c2 = ctx;
  for (const auto& n : f.args) c2.constants.erase(n);
  f.body = expr::parse(text.substr(colon + 1), c2);
  return f;
}

}  // namespace

std::vector<ReduceKey> read_reduce(const ParameterTree& cfg, const std::string& prefix, const expr::Context& ctx) {
  std::vector<ReduceKey> out;
  const std::string p0 = prefix.empty() ? std::string() : prefix + ".";
  for (const auto& name : cfg.sub_keys(prefix)) {
    const std::string p = p0 + name + ".";
    ReduceKey k;
    k.name = name;
    if (cfg.has_key(p + "evaluation.expression")) k.evaluation = expr::parse(cfg.at(p + "evaluation.expression"), ctx);
    if (cfg.has_key(p + "reduction.expression"))
      k.reduction = parse_function(p + "reduction", cfg.at(p + "reduction.expression"), 2, ctx);
    k.initial = cfg.get(p + "initial.value", 0.0);
    if (cfg.has_key(p + "transformation.expression"))
      k.transformation = parse_function(p + "transformation", cfg.at(p + "transformation.expression"), 1, ctx);
    if (cfg.has_key(p + "error.expression")) k.error = parse_function(p + "error", cfg.at(p + "error.expression"), 1, ctx);
    if (cfg.has_key(p + "warn.expression")) k.warn = parse_function(p + "warn", cfg.at(p + "warn.expression"), 1, ctx);
    k.quiet = cfg.get(p + "quiet", false);
    out.push_back(std::move(k));
  }
  return out;
}

double reduce_call(const ReduceFunction& f, const std::vector<double>& x, const std::map<std::string, double>& params) {
  if (!f.body) return 0.0;
  std::map<std::string, double> v = params;
  for (size_t i = 0; i < f.args.size() && i < x.size(); ++i) v[f.args[i]] = x[i];
  return expr::evaluate(f.body, v);
}

std::vector<ReduceKey> Model::reduce_keys(const std::string& reduce_ini) const {
  if (reduce_ini.empty()) return read_reduce(cfg_tree_, "model.reduce", spec_.ctx);
  return read_reduce(ParameterTree::from_ini(reduce_ini, ""), "", spec_.ctx);
}

std::string Model::reduce_source(const std::string& reduce_ini) const {
  std::vector<std::string> kernels;
  return generate_reduce(spec_, L_, reduce_keys(reduce_ini), &kernels);
}

std::vector<ReduceResult> Model::reduce(const std::string& reduce_ini) {
  require_device();
  const std::vector<ReduceKey> keys = reduce_keys(reduce_ini);
  std::vector<ReduceResult> res;
  if (keys.empty()) return res;
  std::vector<std::string> kernels;
  const std::string src = generate_reduce(spec_, L_, keys, &kernels);
  auto mod = rtc_module(src, "dcopasi_reduce", device_);
  std::map<std::string, double> params;
  for (size_t i = 0; i < spec_.params.size(); ++i) params["$" + spec_.params[i]] = spec_.param_values[i];
  const int nk = static_cast<int>(keys.size());
  // one partial (value, defined) per key and workgroup, compartments in order
  std::vector<int> first(L_.comps.size() + 1, 0);
  for (size_t c = 0; c < L_.comps.size(); ++c) {
    const long n = kernels[c].empty() ? 0 : static_cast<long>(L_.comps[c].cells.size());
    first[c + 1] = first[c] + static_cast<int>((n + 255) / 256);
  }
  const int nblk = first.back();
  DeviceArray<double> part;
  part.resize(std::max<size_t>(1, static_cast<size_t>(nblk) * nk * 2));
  const Grid& g = L_.grid;
  int n0 = static_cast<int>(g.n[0]), n1 = g.dim > 1 ? static_cast<int>(g.n[1]) : 1;
  double ox = g.origin[0], oy = g.origin[1], oz = g.origin[2];
  double hx = g.h[0], hy = g.dim > 1 ? g.h[1] : 1.0, hz = g.dim > 2 ? g.h[2] : 1.0;
  double t = time_;
  const double* x = u_.get();
  const double* prm = prm_.get();
  std::vector<DeviceArray<double>> rgeo(L_.comps.size());
  for (size_t c = 0; c < L_.comps.size(); ++c) {
    if (kernels[c].empty() || first[c + 1] == first[c]) continue;
    hipFunction_t fn = mod->function(kernels[c]);
    const int* cells = comp_dev_[c].cells.get();
    const unsigned char* cmask = comp_dev_[c].cell_mask.get();
    const int* cslots = comp_dev_[c].cell_slots.get();
    int ncell = static_cast<int>(L_.comps[c].cells.size());
    double* out = part.get() + static_cast<size_t>(first[c]) * nk * 2;
    // cells of file grids: geometry words at the order-4 points
    DeviceArray<double>& rg = rgeo[c];
    if (g.unstructured) {
      const Rule rule = reduce_rule(g.dim, false);
      rg.upload(cell_geometry(g, L_.comps[c].cells, &rule), stream_);
    }
    const double* rgp = rg.get();
    void* args[] = {&cells, &cmask, &cslots, &ncell, &x, &t, &n0, &n1, &ox, &oy, &oz, &hx, &hy, &hz, &prm, &out, &rgp};
    DCG_HIP(hipModuleLaunchKernel(fn, static_cast<unsigned>(first[c + 1] - first[c]), 1, 1, 256, 1, 1, 0, stream_,
                                  args, nullptr));
  }
  std::vector<double> host(static_cast<size_t>(nblk) * nk * 2);
  if (!host.empty())
    DCG_HIP(hipMemcpyAsync(host.data(), part.get(), host.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
  synchronize();
  for (int k = 0; k < nk; ++k) {
    const auto& K = keys[k];
    auto red = [&](double a, double b) {
      return K.reduction.args.empty() ? a + b : reduce_call(K.reduction, {a, b}, params);
    };
    // partials in cell order, each folded in as the later value
    bool have = false;
    double T = 0.0;
    for (int b = 0; b < nblk; ++b) {
      if (host[(static_cast<size_t>(b) * nk + k) * 2 + 1] == 0.0) continue;
      const double v = host[(static_cast<size_t>(b) * nk + k) * 2];
      T = have ? red(v, T) : v;
      have = true;
    }
    // the reference starts the fold at the initial value and combines the
    // (single, sequential) thread result with it once more (reduce.hh:212-217)
    const double acc = have ? red(T, K.initial) : K.initial;
    ReduceResult r;
    r.name = K.name;
    r.value = red(acc, K.initial);
    if (!K.transformation.args.empty()) r.value = reduce_call(K.transformation, {r.value}, params);
    auto fires = [&](const ReduceFunction& f) {
      // FloatCmp::ne(f(value), 0): any non-zero result
      return !f.args.empty() && reduce_call(f, {r.value}, params) != 0.0;
    };
    if (fires(K.error)) {
      r.flag = 2;
      std::ostringstream m;
      m.precision(17);
      m << "Reduction on the token '" << K.name << "' raised an error: '" << K.error.args[0] << " := " << r.value
        << "'";
      r.message = m.str();
    } else if (fires(K.warn)) {
      r.flag = 1;
    }
    res.push_back(r);
  }
  return res;
}

}  // namespace dcopasi
