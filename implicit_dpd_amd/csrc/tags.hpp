// Tag-value input files of the reference (run_param.txt, config_param.txt):
// ReadTagValues / Matching* of parameter.hpp:16-30,300-376 (';' comments,
// head/tail blanks stripped, later tags override earlier ones, floating
// values rounded through std::stof as at :339, integers through std::stoi).
#pragma once
#include <cstring>
#include <istream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace mdpd {

using TagMap = std::map<std::string, std::vector<std::string>>;

inline const char* const kBlanks = " \t\xE3\x80";  // ASCII blanks and the bytes of U+3000

inline void strip(std::string& b) {
  while (!b.empty() && std::strchr(kBlanks, b.front()) && b.front() != '\0') b.erase(b.begin());
  while (!b.empty() && std::strchr(kBlanks, b.back()) && b.back() != '\0') b.pop_back();
}

inline void read_tags(std::istream& in, TagMap& tv) {
  std::string line, tag;
  while (std::getline(in, line)) {
    const size_t sc = line.find(';');
    if (sc != std::string::npos) line = line.substr(0, sc);
    if (line.empty()) continue;
    strip(line);
    std::stringstream ss(line);
    ss >> tag;
    std::vector<std::string> vals;
    while (!ss.eof()) {
      std::string b;
      ss >> b;
      vals.push_back(b);
    }
    tv[tag] = vals;
  }
}

struct Parse {
  TagMap& tv;
  std::string err;
  bool need(const std::string& tag, size_t num) {
    auto it = tv.find(tag);
    if (it == tv.end()) {
      if (err.empty()) err = "Unmatching occurs. tag: " + tag;
      return false;
    }
    if (num && it->second.size() != num) {
      if (err.empty()) err = "tag " + tag + " needs " + std::to_string(num) + " values";
      return false;
    }
    return true;
  }
  bool f64(const std::string& tag, double* v, size_t num) {
    if (!need(tag, num)) return false;
    try {
      for (size_t i = 0; i < num; ++i) v[i] = (double)std::stof(tv[tag].at(i));
    } catch (...) {
      err = "bad number in tag " + tag;
      return false;
    }
    return true;
  }
  bool u32(const std::string& tag, uint32_t* v, size_t num) {
    if (!need(tag, num)) return false;
    try {
      for (size_t i = 0; i < num; ++i) v[i] = (uint32_t)std::stoi(tv[tag].at(i));
    } catch (...) {
      err = "bad integer in tag " + tag;
      return false;
    }
    return true;
  }
};


}  // namespace mdpd
