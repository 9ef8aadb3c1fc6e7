// hit.h -- reader for the subset of MOOSE's HIT input syntax the driver needs.
//
//   [Block]                 nested blocks, closed by [] or [../]
//     [sub]  key = value    values: bare token, 'quoted list' or "quoted list"
//     []
//   []                      '#' starts a comment outside quotes
//
// Command-line overrides use MOOSE's form Block/sub/key=value.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace mcwhost {

struct HitNode {
  std::string name;                        // last path component
  std::string path;                        // full path, e.g. UserObjects/study
  std::map<std::string, std::string> params;
  std::map<std::string, int> lines;        // line of each parameter
  std::vector<std::unique_ptr<HitNode>> children;
  int line = 0;

  HitNode *child(const std::string &n) const;
  HitNode *find(const std::string &path);  // relative path a/b/c, created on demand = false
  bool has(const std::string &k) const { return params.count(k) != 0; }
  const std::string &get(const std::string &k) const;
};

// Parse text; throws std::runtime_error with "file:line: message".
std::unique_ptr<HitNode> parse_hit(const std::string &text, const std::string &file);

// Apply an override "A/b/key=value" (creating blocks as needed).
void apply_override(HitNode &root, const std::string &assignment);

std::vector<std::string> split_words(const std::string &s);

} // namespace mcwhost
