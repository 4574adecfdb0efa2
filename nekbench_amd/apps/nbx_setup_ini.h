// Reader for nekBone's setup files, with the rules of setupAide
// (core/setupAide.cpp:83-243):
//   [KEY] opens an entry; the text up to the next [ is its value;
//   whitespace outside quotes is dropped from values (keys keep theirs);
//   '#' comments run to the end of the line, /* ... */ comments are skipped;
//   quoted text is kept verbatim, quotes included; \[ and \] are literal.
// get() returns the first entry with the key; compare() is 1 for an equal
// value, 2 when the token occurs inside the value (so "PCG+FLEXIBLE"
// matches "FLEXIBLE" and "NATIVE+SERIAL" matches "SERIAL"), 0 otherwise.
#pragma once
#include <cctype>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

namespace nbx_app {

class SetupIni {
 public:
  bool load(const std::string& path)
  {
    std::ifstream f(path);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    parse(ss.str());
    return true;
  }

  void parse(const std::string& s)
  {
    entries_.clear();
    std::string cur;
    bool have_key = false;
    std::string key;
    auto close_entry = [&]() {
      if (have_key) entries_.emplace_back(key, cur);
      cur.clear();
    };
    const size_t n = s.size();
    for (size_t i = 0; i < n; ++i) {
      const char c = s[i];
      if (c == '\'' || c == '"') {
        size_t j = i + 1;
        while (j < n && s[j] != c) ++j;
        cur.append(s, i, (j < n ? j + 1 : n) - i);
        i = j;
      } else if (c == '/' && i + 1 < n && s[i + 1] == '*') {
        size_t j = s.find("*/", i + 2);
        i = (j == std::string::npos) ? n : j + 1;
      } else if (c == '#') {
        while (i < n && s[i] != '\n') ++i;
      } else if (c == '\\' && i + 1 < n && (s[i + 1] == '[' || s[i + 1] == ']')) {
        cur += s[i + 1];
        ++i;
      } else if (c == '[') {
        close_entry();
        size_t j = i + 1;
        while (j < n && s[j] != ']') ++j;
        key = s.substr(i + 1, j - i - 1);
        have_key = true;
        i = j;
      } else if (!std::isspace(static_cast<unsigned char>(c))) {
        cur += c;
      }
    }
    close_entry();
  }

  bool has(const std::string& k) const
  {
    for (const auto& e : entries_)
      if (e.first == k) return true;
    return false;
  }

  std::string get(const std::string& k) const
  {
    for (const auto& e : entries_)
      if (e.first == k) return e.second;
    return "";
  }

  template <typename T>
  bool get(const std::string& k, T& out) const
  {
    if (!has(k)) return false;
    std::istringstream is(get(k));
    T v;
    if (!(is >> v)) return false;
    out = v;
    return true;
  }

  int compare(const std::string& k, const std::string& token) const
  {
    if (!has(k)) return 0;
    const std::string v = get(k);
    if (v == token) return 1;
    if (v.find(token) != std::string::npos) return 2;
    return 0;
  }

  void set(const std::string& k, const std::string& v)
  {
    for (auto& e : entries_)
      if (e.first == k) {
        e.second = v;
        return;
      }
    entries_.emplace_back(k, v);
  }

  const std::vector<std::pair<std::string, std::string>>& entries() const { return entries_; }

 private:
  std::vector<std::pair<std::string, std::string>> entries_;
};

}  // namespace nbx_app
