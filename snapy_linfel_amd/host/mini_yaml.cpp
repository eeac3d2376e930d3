#include "mini_yaml.hpp"

#include <algorithm>
#include <cctype>
#include <cstring>
#include <fstream>
#include <sstream>

namespace snap {
namespace yaml {

Node Node::operator[](const std::string& key) const {
  if (kind_ != Map) return Node();
  for (auto const& kv : *children_)
    if (kv.first == key) return kv.second;
  return Node();
}

Node Node::operator[](size_t i) const {
  if (kind_ != Seq || i >= items_->size()) return Node();
  return (*items_)[i];
}

size_t Node::size() const {
  if (kind_ == Map) return children_->size();
  if (kind_ == Seq) return items_->size();
  return 0;
}

std::vector<std::string> Node::keys() const {
  std::vector<std::string> out;
  if (kind_ == Map)
    for (auto const& kv : *children_) out.push_back(kv.first);
  return out;
}

void Node::set(const std::string& key, Node v) {
  for (auto& kv : *children_)
    if (kv.first == key) {
      kv.second = std::move(v);
      return;
    }
  children_->emplace_back(key, std::move(v));
}

void Node::push(Node v) { items_->push_back(std::move(v)); }

template <>
std::string Node::as<std::string>() const {
  if (kind_ != Scalar) throw std::runtime_error("yaml: node is not a scalar");
  return value_;
}

template <>
double Node::as<double>() const {
  std::string s = as<std::string>();
  size_t pos = 0;
  double v = std::stod(s, &pos);
  if (pos != s.size()) throw std::runtime_error("yaml: not a number: " + s);
  return v;
}

template <>
int Node::as<int>() const {
  std::string s = as<std::string>();
  size_t pos = 0;
  int v = std::stoi(s, &pos);
  if (pos != s.size()) throw std::runtime_error("yaml: not an integer: " + s);
  return v;
}

template <>
bool Node::as<bool>() const {
  std::string s = as<std::string>();
  std::transform(s.begin(), s.end(), s.begin(), ::tolower);
  if (s == "true" || s == "yes" || s == "on") return true;
  if (s == "false" || s == "no" || s == "off") return false;
  throw std::runtime_error("yaml: not a boolean: " + s);
}

namespace {

struct Line {
  int indent;
  std::string text;  // without indentation and comment
};

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r");
  return s.substr(a, b - a + 1);
}

std::string strip_comment(const std::string& s) {
  char quote = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    char c = s[i];
    if (quote) {
      if (c == quote) quote = 0;
    } else if (c == '"' || c == '\'') {
      quote = c;
    } else if (c == '#' && (i == 0 || isspace((unsigned char)s[i - 1]))) {
      return s.substr(0, i);
    }
  }
  return s;
}

std::string unquote(const std::string& s) {
  if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') ||
                        (s.front() == '\'' && s.back() == '\'')))
    return s.substr(1, s.size() - 2);
  return s;
}

// position of the ':' that separates a key from its value (outside quotes and
// brackets, followed by space or end), or npos
size_t key_colon(const std::string& s) {
  char quote = 0;
  int depth = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    char c = s[i];
    if (quote) {
      if (c == quote) quote = 0;
      continue;
    }
    if (c == '"' || c == '\'') quote = c;
    else if (c == '{' || c == '[') ++depth;
    else if (c == '}' || c == ']') --depth;
    else if (c == ':' && depth == 0 && (i + 1 == s.size() || s[i + 1] == ' ')) return i;
  }
  return std::string::npos;
}

class FlowParser {
 public:
  explicit FlowParser(const std::string& s) : s_(s) {}
  Node parse() {
    Node n = value(",}]");
    skip();
    if (p_ != s_.size()) throw std::runtime_error("yaml: trailing text in flow value: " + s_);
    return n;
  }

 private:
  void skip() {
    while (p_ < s_.size() && isspace((unsigned char)s_[p_])) ++p_;
  }
  Node value(const char* stops) {
    skip();
    if (p_ < s_.size() && s_[p_] == '{') return map();
    if (p_ < s_.size() && s_[p_] == '[') return seq();
    return Node::scalar(unquote(trim(token(stops))));
  }
  std::string token(const char* stops) {
    size_t start = p_;
    char quote = 0;
    while (p_ < s_.size()) {
      char c = s_[p_];
      if (quote) {
        if (c == quote) quote = 0;
      } else if (c == '"' || c == '\'') {
        quote = c;
      } else if (strchr(stops, c)) {
        break;
      }
      ++p_;
    }
    return s_.substr(start, p_ - start);
  }
  Node map() {
    Node m = Node::map();
    ++p_;  // {
    for (;;) {
      skip();
      if (p_ < s_.size() && s_[p_] == '}') {
        ++p_;
        return m;
      }
      std::string key = unquote(trim(token(":,}")));
      if (p_ >= s_.size() || s_[p_] != ':') throw std::runtime_error("yaml: bad flow map: " + s_);
      ++p_;
      m.set(key, value(",}"));
      skip();
      if (p_ < s_.size() && s_[p_] == ',') ++p_;
    }
  }
  Node seq() {
    Node q = Node::seq();
    ++p_;  // [
    for (;;) {
      skip();
      if (p_ < s_.size() && s_[p_] == ']') {
        ++p_;
        return q;
      }
      q.push(value(",]"));
      skip();
      if (p_ < s_.size() && s_[p_] == ',') ++p_;
      if (p_ >= s_.size()) throw std::runtime_error("yaml: unterminated flow sequence: " + s_);
    }
  }
  const std::string& s_;
  size_t p_ = 0;
};

Node inline_value(const std::string& text) {
  std::string t = trim(text);
  if (t.empty()) return Node();
  return FlowParser(t).parse();
}

class BlockParser {
 public:
  explicit BlockParser(std::vector<Line> lines) : L_(std::move(lines)) {}
  Node parse() {
    if (L_.empty()) return Node::map();
    return block(L_[0].indent);
  }

 private:
  Node block(int indent) {
    if (i_ < L_.size() && L_[i_].text.rfind("-", 0) == 0 &&
        (L_[i_].text.size() == 1 || L_[i_].text[1] == ' '))
      return sequence(indent);
    return mapping(indent);
  }

  Node mapping(int indent) {
    Node m = Node::map();
    while (i_ < L_.size() && L_[i_].indent == indent && L_[i_].text.rfind("- ", 0) != 0 &&
           L_[i_].text != "-") {
      const std::string& t = L_[i_].text;
      size_t c = key_colon(t);
      if (c == std::string::npos) throw std::runtime_error("yaml: expected 'key: value': " + t);
      std::string key = unquote(trim(t.substr(0, c)));
      std::string rest = trim(t.substr(c + 1));
      ++i_;
      m.set(key, rest.empty() ? nested(indent) : inline_value(rest));
    }
    return m;
  }

  // value of a key whose text is empty: a deeper block, a same-indent
  // sequence, or null
  Node nested(int indent) {
    if (i_ >= L_.size()) return Node();
    if (L_[i_].indent > indent) return block(L_[i_].indent);
    if (L_[i_].indent == indent && L_[i_].text.rfind("- ", 0) == 0) return sequence(indent);
    return Node();
  }

  Node sequence(int indent) {
    Node q = Node::seq();
    while (i_ < L_.size() && L_[i_].indent == indent &&
           (L_[i_].text.rfind("- ", 0) == 0 || L_[i_].text == "-")) {
      std::string rest = L_[i_].text == "-" ? "" : trim(L_[i_].text.substr(2));
      if (rest.empty()) {
        ++i_;
        q.push(nested(indent));
      } else if (key_colon(rest) != std::string::npos && rest[0] != '{' && rest[0] != '[') {
        // "- key: value" starts a mapping whose further keys sit at the
        // column of "key"
        int inner = indent + (int)L_[i_].text.find_first_not_of(' ', 1);
        L_[i_].indent = inner;
        L_[i_].text = rest;
        q.push(mapping(inner));
      } else {
        ++i_;
        q.push(inline_value(rest));
      }
    }
    return q;
  }

  std::vector<Line> L_;
  size_t i_ = 0;
};

}  // namespace

Node Load(const std::string& text) {
  std::vector<Line> lines;
  std::istringstream in(text);
  std::string raw;
  while (std::getline(in, raw)) {
    std::string s = strip_comment(raw);
    if (trim(s).empty()) continue;
    if (trim(s) == "---" || trim(s) == "...") continue;
    int ind = 0;
    while (ind < (int)s.size() && s[ind] == ' ') ++ind;
    lines.push_back({ind, trim(s)});
  }
  return BlockParser(std::move(lines)).parse();
}

Node LoadFile(const std::string& filename) {
  std::ifstream f(filename);
  if (!f) throw std::runtime_error("yaml: cannot open " + filename);
  std::stringstream ss;
  ss << f.rdbuf();
  return Load(ss.str());
}

}  // namespace yaml
}  // namespace snap
