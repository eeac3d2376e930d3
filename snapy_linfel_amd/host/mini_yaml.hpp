// Minimal YAML reader for the hot-path configuration files.
//
// Supports what the reference's input files use: block mappings by
// indentation, block sequences ("- "), flow mappings ({k: v, ...}) and flow
// sequences ([a, b]), plain / quoted scalars and # comments.  Lookups of
// missing keys return a null node, and as<T>(fallback) returns the fallback
// for null nodes, mirroring yaml-cpp's node["key"].as<T>(default) idiom.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace snap {
namespace yaml {

class Node {
 public:
  enum Kind { Null, Scalar, Map, Seq };

  Node() = default;
  static Node scalar(std::string s) {
    Node n;
    n.kind_ = Scalar;
    n.value_ = std::move(s);
    return n;
  }
  static Node map() {
    Node n;
    n.kind_ = Map;
    n.children_ = std::make_shared<std::vector<std::pair<std::string, Node>>>();
    return n;
  }
  static Node seq() {
    Node n;
    n.kind_ = Seq;
    n.items_ = std::make_shared<std::vector<Node>>();
    return n;
  }

  Kind kind() const { return kind_; }
  bool defined() const { return kind_ != Null; }
  explicit operator bool() const { return defined(); }
  bool is_map() const { return kind_ == Map; }
  bool is_seq() const { return kind_ == Seq; }
  bool is_scalar() const { return kind_ == Scalar; }
  const std::string& str() const { return value_; }

  Node operator[](const std::string& key) const;
  Node operator[](size_t i) const;
  size_t size() const;
  std::vector<std::string> keys() const;

  void set(const std::string& key, Node v);
  void push(Node v);

  template <typename T>
  T as() const;
  template <typename T>
  T as(const T& fallback) const {
    return defined() ? as<T>() : fallback;
  }

 private:
  Kind kind_ = Null;
  std::string value_;
  std::shared_ptr<std::vector<std::pair<std::string, Node>>> children_;
  std::shared_ptr<std::vector<Node>> items_;
};

template <>
std::string Node::as<std::string>() const;
template <>
double Node::as<double>() const;
template <>
int Node::as<int>() const;
template <>
bool Node::as<bool>() const;

Node Load(const std::string& text);
Node LoadFile(const std::string& filename);

}  // namespace yaml
}  // namespace snap
