// json_min.hpp — reader for the flat JSON objects of sim_config.json and
// pigment_config.json ({"key": number, ...}), the subset
// Simulation::setupScene reads (simulation_units.cpp:32-80).  Booleans map to
// 1/0; nested values are skipped.
#pragma once
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>

namespace mpmx {

class JsonFlat {
 public:
  bool load(const char* path) {
    std::ifstream f(path);
    if (!f.is_open()) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
  }
  bool has(const char* k) const { return vals_.count(k) != 0; }
  float num(const char* k, float def) const {
    auto it = vals_.find(k);
    return it == vals_.end() ? def : (float)it->second;
  }

 private:
  std::map<std::string, double> vals_;
  static void ws(const std::string& s, size_t& i) {
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
  }
  static bool str(const std::string& s, size_t& i, std::string& out) {
    if (i >= s.size() || s[i] != '"') return false;
    ++i;
    out.clear();
    while (i < s.size() && s[i] != '"') {
      if (s[i] == '\\' && i + 1 < s.size()) ++i;
      out.push_back(s[i++]);
    }
    if (i >= s.size()) return false;
    ++i;
    return true;
  }
  // skip any JSON value (for nested objects/arrays/strings)
  static bool skip(const std::string& s, size_t& i) {
    ws(s, i);
    if (i >= s.size()) return false;
    if (s[i] == '"') { std::string t; return str(s, i, t); }
    if (s[i] == '{' || s[i] == '[') {
      int depth = 0;
      do {
        if (s[i] == '"') { std::string t; if (!str(s, i, t)) return false; continue; }
        if (s[i] == '{' || s[i] == '[') ++depth;
        if (s[i] == '}' || s[i] == ']') --depth;
        ++i;
      } while (i < s.size() && depth > 0);
      return depth == 0;
    }
    while (i < s.size() && s[i] != ',' && s[i] != '}') ++i;
    return true;
  }
  bool parse(const std::string& s) {
    size_t i = 0;
    ws(s, i);
    if (i >= s.size() || s[i] != '{') return false;
    ++i;
    for (;;) {
      ws(s, i);
      if (i < s.size() && s[i] == '}') return true;
      std::string key;
      if (!str(s, i, key)) return false;
      ws(s, i);
      if (i >= s.size() || s[i] != ':') return false;
      ++i;
      ws(s, i);
      if (i >= s.size()) return false;
      if (s.compare(i, 4, "true") == 0) { vals_[key] = 1.0; i += 4; }
      else if (s.compare(i, 5, "false") == 0) { vals_[key] = 0.0; i += 5; }
      else if (s[i] == '-' || s[i] == '+' || s[i] == '.' || std::isdigit((unsigned char)s[i])) {
        const char* b = s.c_str() + i;
        char* e = nullptr;
        double v = std::strtod(b, &e);
        if (e == b) return false;
        vals_[key] = v;
        i += (size_t)(e - b);
      } else if (!skip(s, i)) {
        return false;
      }
      ws(s, i);
      if (i < s.size() && s[i] == ',') { ++i; continue; }
      if (i < s.size() && s[i] == '}') return true;
      return false;
    }
  }
};

}  // namespace mpmx
