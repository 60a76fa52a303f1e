// Parameter tree read from DuneCopasi INI text.
//
// Same key semantics as Dune::ParameterTree as used by the reference CLI
// (src/dune_copasi.cc:270-288): "[section]" prefixes, dotted keys,
// "--key=value" style overrides applied last, and sub-keys enumerated in the
// order they first appear (the reference orders compartments and species that
// way, model/diffusion_reaction/model_multi_compartment.impl.hh:63-75).
#pragma once
#include <string>
#include <unordered_map>
#include <vector>

namespace dcopasi {

class ParameterTree {
 public:
  static ParameterTree from_ini(const std::string& text, const std::string& overrides);

  bool has_key(const std::string& key) const { return map_.count(key) != 0; }
  bool has_sub(const std::string& prefix) const;
  const std::string& at(const std::string& key) const;
  std::string get(const std::string& key, const std::string& fallback) const;
  double get(const std::string& key, double fallback) const;
  int get(const std::string& key, int fallback) const;
  bool get(const std::string& key, bool fallback) const;
  std::vector<double> get_list(const std::string& key) const;
  std::vector<std::string> sub_keys(const std::string& prefix) const;
  void set(const std::string& key, const std::string& value);

 private:
  void read(const std::string& text, bool overrides);
  std::unordered_map<std::string, std::string> map_;
  std::vector<std::string> keys_;  // first-appearance order
};

}  // namespace dcopasi
