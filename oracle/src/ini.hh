// Oracle INI reader: a plain restatement of the Dune::ParameterTree INI
// semantics that DuneCopasi reads its configuration with
// (src/dune_copasi.cc:270-288): "[a.b]" section prefixes, "key = value"
// assignments, '#' comments, later assignments overriding earlier ones, and
// sub-keys listed in first-appearance order (compartment and species order in
// model/diffusion_reaction/model_multi_compartment.impl.hh:63-75 follows it).
//
// Test infrastructure only: used by the oracle, never by the product.
#pragma once
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

inline std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

class Ini {
 public:
  void parse(const std::string& text) {
    std::istringstream in(text);
    std::string line, section;
    while (std::getline(in, line)) {
      // strip comments (no quoting support needed for the keys we read)
      auto h = line.find('#');
      if (h != std::string::npos) line = line.substr(0, h);
      line = trim(line);
      if (line.empty()) continue;
      if (line.front() == '[') {
        auto e = line.find(']');
        if (e == std::string::npos) throw std::runtime_error("ini: bad section line: " + line);
        section = trim(line.substr(1, e - 1));
        continue;
      }
      auto eq = line.find('=');
      if (eq == std::string::npos) throw std::runtime_error("ini: missing '=' in: " + line);
      std::string key = trim(line.substr(0, eq));
      std::string val = trim(line.substr(eq + 1));
      if (val.size() >= 2 && val.front() == '"' && val.back() == '"') val = val.substr(1, val.size() - 2);
      set(section.empty() ? key : section + "." + key, val);
    }
  }
  // "key=value" lines (the CLI's --key=value overrides)
  void parse_overrides(const std::string& text) {
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
      line = trim(line);
      if (line.empty()) continue;
      auto eq = line.find('=');
      if (eq == std::string::npos) throw std::runtime_error("ini: bad override: " + line);
      set(trim(line.substr(0, eq)), trim(line.substr(eq + 1)));
    }
  }
  void set(const std::string& k, const std::string& v) {
    if (!values_.count(k)) order_.push_back(k);
    values_[k] = v;
  }
  bool has(const std::string& k) const { return values_.count(k) != 0; }
  std::string get(const std::string& k, const std::string& def = "") const {
    auto it = values_.find(k);
    return it == values_.end() ? def : it->second;
  }
  double get_double(const std::string& k, double def) const {
    auto it = values_.find(k);
    if (it == values_.end()) return def;
    return std::stod(it->second);
  }
  std::vector<double> get_doubles(const std::string& k) const {
    std::vector<double> r;
    std::istringstream s(get(k));
    double x;
    while (s >> x) r.push_back(x);
    return r;
  }
  bool has_sub(const std::string& prefix) const {
    std::string p = prefix + ".";
    for (auto& k : order_)
      if (k.compare(0, p.size(), p) == 0) return true;
    return false;
  }
  // immediate children of `prefix` in first-appearance order
  std::vector<std::string> subkeys(const std::string& prefix) const {
    std::vector<std::string> r;
    std::string p = prefix.empty() ? "" : prefix + ".";
    for (auto& k : order_) {
      if (k.compare(0, p.size(), p) != 0) continue;
      std::string rest = k.substr(p.size());
      auto d = rest.find('.');
      if (d == std::string::npos) continue;  // a value key, not a sub-tree
      std::string c = rest.substr(0, d);
      bool seen = false;
      for (auto& x : r) seen |= (x == c);
      if (!seen) r.push_back(c);
    }
    return r;
  }

 private:
  std::map<std::string, std::string> values_;
  std::vector<std::string> order_;
};

}  // namespace orc
