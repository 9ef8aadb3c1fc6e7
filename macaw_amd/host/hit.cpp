// hit.cpp -- HIT-subset reader (see hit.h).
#include "hit.h"

#include <sstream>
#include <stdexcept>

namespace mcwhost {

HitNode *HitNode::child(const std::string &n) const
{
  for (auto &c : children)
    if (c->name == n) return c.get();
  return nullptr;
}

HitNode *HitNode::find(const std::string &p)
{
  HitNode *cur = this;
  std::string rest = p;
  while (!rest.empty() && cur) {
    auto k = rest.find('/');
    std::string head = rest.substr(0, k);
    rest = k == std::string::npos ? "" : rest.substr(k + 1);
    if (head.empty()) continue;
    cur = cur->child(head);
  }
  return cur;
}

const std::string &HitNode::get(const std::string &k) const
{
  auto it = params.find(k);
  if (it == params.end()) throw std::runtime_error(path + ": missing required parameter '" + k + "'");
  return it->second;
}

std::vector<std::string> split_words(const std::string &s)
{
  std::vector<std::string> out;
  std::istringstream is(s);
  std::string w;
  while (is >> w) out.push_back(w);
  return out;
}

static std::string trim(const std::string &s)
{
  size_t a = s.find_first_not_of(" \t\r");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r");
  return s.substr(a, b - a + 1);
}

static HitNode *ensure(HitNode *parent, const std::string &name, int line)
{
  // a header like [Variables/temperature] opens nested blocks
  HitNode *cur = parent;
  std::string rest = name;
  while (!rest.empty()) {
    auto k = rest.find('/');
    std::string head = rest.substr(0, k);
    rest = k == std::string::npos ? "" : rest.substr(k + 1);
    HitNode *c = cur->child(head);
    if (!c) {
      auto n = std::make_unique<HitNode>();
      n->name = head;
      n->path = cur->path.empty() ? head : cur->path + "/" + head;
      n->line = line;
      c = n.get();
      cur->children.push_back(std::move(n));
    }
    cur = c;
  }
  return cur;
}

std::unique_ptr<HitNode> parse_hit(const std::string &text, const std::string &file)
{
  auto root = std::make_unique<HitNode>();
  std::vector<std::pair<HitNode *, HitNode *>> stack; // (opened node, node to return to)
  HitNode *cur = root.get();
  std::istringstream is(text);
  std::string raw;
  int lineno = 0;
  auto fail = [&](const std::string &m) { throw std::runtime_error(file + ":" + std::to_string(lineno) + ": " + m); };
  std::string pending; // multi-line quoted value continuation
  std::string pending_key;
  char pending_q = 0;
  while (std::getline(is, raw)) {
    ++lineno;
    if (pending_q) {
      auto q = raw.find(pending_q);
      if (q == std::string::npos) {
        pending += " " + raw;
        continue;
      }
      pending += " " + raw.substr(0, q);
      cur->params[pending_key] = trim(pending);
      pending_q = 0;
      raw = raw.substr(q + 1);
    }
    // strip comments outside quotes
    std::string line;
    char q = 0;
    for (char ch : raw) {
      if (q) {
        if (ch == q) q = 0;
      } else if (ch == '\'' || ch == '"') {
        q = ch;
      } else if (ch == '#') {
        break;
      }
      line += ch;
    }
    line = trim(line);
    while (!line.empty()) {
      if (line[0] == '[') {
        auto e = line.find(']');
        if (e == std::string::npos) fail("unterminated block header");
        std::string name = trim(line.substr(1, e - 1));
        line = trim(line.substr(e + 1));
        if (name.empty() || name == "../" || name == "..") {
          if (stack.empty()) fail("unmatched block close");
          cur = stack.back().second;
          stack.pop_back();
        } else {
          if (name.rfind("./", 0) == 0) name = name.substr(2);
          HitNode *n = ensure(cur, name, lineno);
          stack.push_back({n, cur});
          cur = n;
        }
        continue;
      }
      auto eq = line.find('=');
      if (eq == std::string::npos) fail("expected key = value");
      std::string key = trim(line.substr(0, eq));
      std::string rest = trim(line.substr(eq + 1));
      std::string val;
      if (!rest.empty() && (rest[0] == '\'' || rest[0] == '"')) {
        char qq = rest[0];
        auto e = rest.find(qq, 1);
        if (e == std::string::npos) {
          pending = rest.substr(1);
          pending_key = key;
          pending_q = qq;
          cur->lines[key] = lineno;
          line.clear();
          break;
        }
        val = rest.substr(1, e - 1);
        line = trim(rest.substr(e + 1));
      } else {
        auto sp = rest.find_first_of(" \t");
        val = rest.substr(0, sp);
        line = sp == std::string::npos ? "" : trim(rest.substr(sp));
      }
      cur->params[key] = trim(val);
      cur->lines[key] = lineno;
    }
  }
  if (pending_q) fail("unterminated quoted value");
  if (!stack.empty()) fail("block [" + stack.back().first->path + "] is not closed");
  return root;
}

void apply_override(HitNode &root, const std::string &a)
{
  auto eq = a.find('=');
  if (eq == std::string::npos) throw std::runtime_error("override must be Path/key=value: " + a);
  std::string lhs = a.substr(0, eq), val = a.substr(eq + 1);
  if (val.size() >= 2 && (val[0] == '\'' || val[0] == '"')) val = val.substr(1, val.size() - 2);
  auto s = lhs.rfind('/');
  std::string blk = s == std::string::npos ? "" : lhs.substr(0, s);
  std::string key = s == std::string::npos ? lhs : lhs.substr(s + 1);
  HitNode *n = blk.empty() ? &root : ensure(&root, blk, 0);
  n->params[key] = val;
}

} // namespace mcwhost
