#include "config.hh"

#include <sstream>

#include "error.hh"

namespace dcopasi {

static std::string strip(const std::string& s) {
  const char* ws = " \t\r\n";
  auto a = s.find_first_not_of(ws);
  if (a == std::string::npos) return {};
  auto b = s.find_last_not_of(ws);
  return s.substr(a, b - a + 1);
}

ParameterTree ParameterTree::from_ini(const std::string& text, const std::string& overrides) {
  ParameterTree t;
  t.read(text, false);
  t.read(overrides, true);
  return t;
}

void ParameterTree::read(const std::string& text, bool overrides) {
  std::istringstream in(text);
  std::string raw, prefix;
  int lineno = 0;
  while (std::getline(in, raw)) {
    ++lineno;
    std::string line = raw;
    if (!overrides) {
      auto hash = line.find('#');
      if (hash != std::string::npos) line.resize(hash);
    }
    line = strip(line);
    if (line.empty()) continue;
    if (!overrides && line[0] == '[') {
      auto close = line.find(']');
      if (close == std::string::npos) throw IOError("INI line " + std::to_string(lineno) + ": unterminated section");
      prefix = strip(line.substr(1, close - 1));
      continue;
    }
    if (overrides && line.rfind("--", 0) == 0) line = line.substr(2);
    auto eq = line.find('=');
    if (eq == std::string::npos) throw IOError("INI line " + std::to_string(lineno) + ": expected 'key = value'");
    std::string key = strip(line.substr(0, eq));
    std::string value = strip(line.substr(eq + 1));
    if (value.size() >= 2 && value.front() == '"' && value.back() == '"') value = value.substr(1, value.size() - 2);
    set(prefix.empty() || overrides ? key : prefix + "." + key, value);
  }
}

void ParameterTree::set(const std::string& key, const std::string& value) {
  auto it = map_.find(key);
  if (it == map_.end()) {
    keys_.push_back(key);
    map_.emplace(key, value);
  } else {
    it->second = value;
  }
}

bool ParameterTree::has_sub(const std::string& prefix) const {
  const std::string p = prefix + ".";
  for (const auto& k : keys_)
    if (k.size() > p.size() && k.compare(0, p.size(), p) == 0) return true;
  return false;
}

const std::string& ParameterTree::at(const std::string& key) const {
  auto it = map_.find(key);
  if (it == map_.end()) throw IOError("missing configuration key '" + key + "'");
  return it->second;
}

std::string ParameterTree::get(const std::string& key, const std::string& fallback) const {
  auto it = map_.find(key);
  return it == map_.end() ? fallback : it->second;
}

double ParameterTree::get(const std::string& key, double fallback) const {
  auto it = map_.find(key);
  if (it == map_.end()) return fallback;
  try {
    return std::stod(it->second);
  } catch (...) {
    throw IOError("key '" + key + "' is not a number: '" + it->second + "'");
  }
}

int ParameterTree::get(const std::string& key, int fallback) const {
  auto it = map_.find(key);
  if (it == map_.end()) return fallback;
  try {
    return std::stoi(it->second);
  } catch (...) {
    throw IOError("key '" + key + "' is not an integer: '" + it->second + "'");
  }
}

bool ParameterTree::get(const std::string& key, bool fallback) const {
  auto it = map_.find(key);
  if (it == map_.end()) return fallback;
  const std::string& v = it->second;
  if (v == "true" || v == "1" || v == "yes") return true;
  if (v == "false" || v == "0" || v == "no") return false;
  throw IOError("key '" + key + "' is not a boolean: '" + v + "'");
}

std::vector<double> ParameterTree::get_list(const std::string& key) const {
  std::vector<double> out;
  auto it = map_.find(key);
  if (it == map_.end()) return out;
  std::istringstream s(it->second);
  double v;
  while (s >> v) out.push_back(v);
  return out;
}

std::vector<std::string> ParameterTree::sub_keys(const std::string& prefix) const {
  std::vector<std::string> out;
  const std::string p = prefix.empty() ? std::string() : prefix + ".";
  for (const auto& k : keys_) {
    if (k.compare(0, p.size(), p) != 0) continue;
    auto dot = k.find('.', p.size());
    if (dot == std::string::npos) continue;
    std::string child = k.substr(p.size(), dot - p.size());
    bool seen = false;
    for (const auto& c : out) seen = seen || c == child;
    if (!seen) out.push_back(child);
  }
  return out;
}

}  // namespace dcopasi
