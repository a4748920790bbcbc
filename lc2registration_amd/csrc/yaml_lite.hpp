// A small YAML subset for the registration configuration surface (parameters*.yaml):
// block mappings, block sequences (including nested "-" items), flow mappings and
// flow sequences, plain / single- / double-quoted scalars, comments. Mapping order and
// flow style are kept so a loaded file is written back in the same shape
// (the reference writes its result file with yaml-cpp, optimizers.cu:1040-1106).
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace lc2yaml {

class Node {
public:
    enum Kind { Null, Scalar, Map, Seq };

    Node() = default;
    static Node scalar(const std::string& s, bool quoted = false);
    static Node number(double v, int digits = 17);
    static Node boolean(bool v);
    static Node map(bool flow = false);
    static Node seq(bool flow = false);

    Kind kind() const { return kind_; }
    bool is_null() const { return kind_ == Null; }
    bool is_map() const { return kind_ == Map; }
    bool is_seq() const { return kind_ == Seq; }
    bool is_scalar() const { return kind_ == Scalar; }
    bool flow() const { return flow_; }
    void set_flow(bool f) { flow_ = f; }
    size_t size() const;

    // mapping access
    bool has(const std::string& key) const;
    const Node& at(const std::string& key) const;  // throws if missing
    Node& operator[](const std::string& key);      // inserts (turns Null into Map)
    const std::vector<std::pair<std::string, Node>>& items() const { return map_; }
    // sequence access
    const Node& at(size_t i) const;
    void push_back(const Node& n);
    const std::vector<Node>& elements() const { return seq_; }

    // scalar conversion (throw std::runtime_error with the offending text)
    const std::string& str() const;
    double as_double() const;
    int as_int() const;
    bool as_bool() const;
    bool quoted() const { return quoted_; }

private:
    Kind kind_ = Null;
    bool flow_ = false;
    bool quoted_ = false;
    std::string scalar_;
    std::vector<std::pair<std::string, Node>> map_;
    std::vector<Node> seq_;
};

Node parse(const std::string& text);
Node load_file(const std::string& path);
std::string dump(const Node& root);
void save_file(const Node& root, const std::string& path);

// Path lookup "a/b/c" (nullptr if absent).
const Node* find(const Node& root, const std::string& path);

}  // namespace lc2yaml
