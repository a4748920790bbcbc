// YAML subset reader/writer (see yaml_lite.hpp).
#include "yaml_lite.hpp"

#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace lc2yaml {

Node Node::scalar(const std::string& s, bool quoted) {
    Node n;
    n.kind_ = Scalar;
    n.scalar_ = s;
    n.quoted_ = quoted;
    return n;
}

Node Node::number(double v, int digits) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.*g", digits, v);
    return scalar(buf);
}

Node Node::boolean(bool v) { return scalar(v ? "true" : "false"); }

Node Node::map(bool flow) {
    Node n;
    n.kind_ = Map;
    n.flow_ = flow;
    return n;
}

Node Node::seq(bool flow) {
    Node n;
    n.kind_ = Seq;
    n.flow_ = flow;
    return n;
}

size_t Node::size() const {
    if (kind_ == Map) return map_.size();
    if (kind_ == Seq) return seq_.size();
    return 0;
}

bool Node::has(const std::string& key) const {
    if (kind_ != Map) return false;
    for (auto& kv : map_)
        if (kv.first == key) return true;
    return false;
}

const Node& Node::at(const std::string& key) const {
    if (kind_ == Map)
        for (auto& kv : map_)
            if (kv.first == key) return kv.second;
    throw std::runtime_error("yaml: missing key '" + key + "'");
}

Node& Node::operator[](const std::string& key) {
    if (kind_ == Null) kind_ = Map;
    if (kind_ != Map) throw std::runtime_error("yaml: '" + key + "' indexed on a non-mapping");
    for (auto& kv : map_)
        if (kv.first == key) return kv.second;
    map_.emplace_back(key, Node());
    return map_.back().second;
}

const Node& Node::at(size_t i) const {
    if (kind_ != Seq || i >= seq_.size()) throw std::runtime_error("yaml: sequence index out of range");
    return seq_[i];
}

void Node::push_back(const Node& n) {
    if (kind_ == Null) kind_ = Seq;
    if (kind_ != Seq) throw std::runtime_error("yaml: push_back on a non-sequence");
    seq_.push_back(n);
}

const std::string& Node::str() const {
    if (kind_ != Scalar) throw std::runtime_error("yaml: expected a scalar");
    return scalar_;
}

double Node::as_double() const {
    const std::string& s = str();
    char* end = nullptr;
    errno = 0;
    double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || *end != '\0' || errno == ERANGE)
        throw std::runtime_error("yaml: '" + s + "' is not a number");
    return v;
}

int Node::as_int() const {
    const std::string& s = str();
    char* end = nullptr;
    long v = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str() || *end != '\0') {
        double d = as_double();
        if (d != (double)(long)d) throw std::runtime_error("yaml: '" + s + "' is not an integer");
        return (int)d;
    }
    return (int)v;
}

bool Node::as_bool() const {
    std::string s = str();
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    if (s == "true" || s == "yes" || s == "on") return true;
    if (s == "false" || s == "no" || s == "off") return false;
    throw std::runtime_error("yaml: '" + str() + "' is not a boolean");
}

namespace {

struct Line {
    int indent;
    std::string text;
    int lineno;
};

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// Remove a trailing comment (a '#' at line start or after whitespace, outside quotes).
std::string strip_comment(const std::string& s) {
    char q = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (q) {
            if (c == '\\' && q == '"') ++i;
            else if (c == q) q = 0;
            continue;
        }
        if (c == '"' || c == '\'') q = c;
        else if (c == '#' && (i == 0 || std::isspace((unsigned char)s[i - 1]))) return s.substr(0, i);
    }
    return s;
}

std::vector<Line> split_lines(const std::string& text) {
    std::vector<Line> out;
    std::istringstream is(text);
    std::string raw;
    int no = 0;
    while (std::getline(is, raw)) {
        ++no;
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        std::string t = strip_comment(raw);
        if (trim(t).empty()) continue;
        if (trim(t) == "---" || trim(t) == "...") continue;
        int ind = 0;
        while (ind < (int)t.size() && t[ind] == ' ') ++ind;
        if (ind < (int)t.size() && t[ind] == '\t') throw std::runtime_error("yaml: tab indentation, line " + std::to_string(no));
        out.push_back({ind, trim(t), no});
    }
    return out;
}

bool is_seq_item(const std::string& t) { return !t.empty() && t[0] == '-' && (t.size() == 1 || t[1] == ' '); }

std::string unquote(const std::string& s, bool* quoted) {
    *quoted = false;
    if (s.size() >= 2 && s.front() == '"' && s.back() == '"') {
        *quoted = true;
        std::string o;
        for (size_t i = 1; i + 1 < s.size(); ++i) {
            char c = s[i];
            if (c == '\\' && i + 2 < s.size()) {
                char e = s[++i];
                switch (e) {
                    case 'n': o += '\n'; break;
                    case 't': o += '\t'; break;
                    case 'r': o += '\r'; break;
                    case '0': o += '\0'; break;
                    default: o += e; break;
                }
            } else {
                o += c;
            }
        }
        return o;
    }
    if (s.size() >= 2 && s.front() == '\'' && s.back() == '\'') {
        *quoted = true;
        std::string o;
        for (size_t i = 1; i + 1 < s.size(); ++i) {
            if (s[i] == '\'' && i + 2 < s.size() && s[i + 1] == '\'') ++i;
            o += s[i];
        }
        return o;
    }
    return s;
}

// Position of the "key: " separator outside quotes and brackets, or npos.
size_t key_sep(const std::string& t) {
    char q = 0;
    int depth = 0;
    for (size_t i = 0; i < t.size(); ++i) {
        const char c = t[i];
        if (q) {
            if (c == '\\' && q == '"') ++i;
            else if (c == q) q = 0;
            continue;
        }
        if (c == '"' || c == '\'') q = c;
        else if (c == '{' || c == '[') ++depth;
        else if (c == '}' || c == ']') --depth;
        else if (c == ':' && depth == 0 && (i + 1 == t.size() || t[i + 1] == ' ')) return i;
    }
    return std::string::npos;
}

struct Flow {
    const std::string& s;
    size_t p = 0;
    int lineno;
    void ws() {
        while (p < s.size() && std::isspace((unsigned char)s[p])) ++p;
    }
    [[noreturn]] void err(const char* what) {
        throw std::runtime_error(std::string("yaml: ") + what + " in flow collection, line " +
                                 std::to_string(lineno));
    }
    std::string token(const char* stops) {
        ws();
        size_t a = p;
        char q = 0;
        if (p < s.size() && (s[p] == '"' || s[p] == '\'')) {
            q = s[p++];
            while (p < s.size() && s[p] != q) {
                if (s[p] == '\\' && q == '"') ++p;
                ++p;
            }
            if (p >= s.size()) err("unterminated quote");
            ++p;
            return s.substr(a, p - a);
        }
        while (p < s.size() && !std::strchr(stops, s[p])) ++p;
        return trim(s.substr(a, p - a));
    }
    Node value() {
        ws();
        if (p >= s.size()) err("unexpected end");
        if (s[p] == '{') {
            ++p;
            Node m = Node::map(true);
            ws();
            if (p < s.size() && s[p] == '}') { ++p; return m; }
            while (true) {
                std::string k = token(":,}");
                if (p >= s.size() || s[p] != ':') err("expected ':'");
                ++p;
                bool q;
                std::string key = unquote(k, &q);
                m[key] = value();
                ws();
                if (p < s.size() && s[p] == ',') { ++p; continue; }
                if (p < s.size() && s[p] == '}') { ++p; break; }
                err("expected ',' or '}'");
            }
            return m;
        }
        if (s[p] == '[') {
            ++p;
            Node q = Node::seq(true);
            ws();
            if (p < s.size() && s[p] == ']') { ++p; return q; }
            while (true) {
                q.push_back(value());
                ws();
                if (p < s.size() && s[p] == ',') { ++p; continue; }
                if (p < s.size() && s[p] == ']') { ++p; break; }
                err("expected ',' or ']'");
            }
            return q;
        }
        bool quoted;
        std::string t = token(",}]");
        std::string v = unquote(t, &quoted);
        if (!quoted && (v == "~" || v == "null")) return Node();
        return Node::scalar(v, quoted);
    }
};

Node inline_value(const std::string& t, int lineno) {
    if (!t.empty() && (t[0] == '{' || t[0] == '[')) {
        Flow f{t, 0, lineno};
        Node n = f.value();
        f.ws();
        if (f.p != t.size()) throw std::runtime_error("yaml: trailing text, line " + std::to_string(lineno));
        return n;
    }
    bool quoted;
    std::string v = unquote(t, &quoted);
    if (!quoted && (v == "~" || v == "null")) return Node();
    return Node::scalar(v, quoted);
}

Node parse_block(std::vector<Line>& L, size_t& i, int indent);

Node parse_seq(std::vector<Line>& L, size_t& i, int indent) {
    Node n = Node::seq(false);
    while (i < L.size() && L[i].indent == indent && is_seq_item(L[i].text)) {
        std::string rest = L[i].text.size() > 1 ? trim(L[i].text.substr(1)) : std::string();
        if (rest.empty()) {
            ++i;
            if (i < L.size() && L[i].indent > indent) n.push_back(parse_block(L, i, L[i].indent));
            else n.push_back(Node());
        } else if (!(rest[0] == '{' || rest[0] == '[' || rest[0] == '"' || rest[0] == '\'') &&
                   (key_sep(rest) != std::string::npos || is_seq_item(rest))) {
            // "- key: value" opens a mapping (or nested sequence) at the item's column
            const int col = indent + (int)(L[i].text.size() - rest.size());
            L[i].indent = col;
            L[i].text = rest;
            n.push_back(parse_block(L, i, col));
        } else {
            n.push_back(inline_value(rest, L[i].lineno));
            ++i;
        }
    }
    return n;
}

Node parse_map(std::vector<Line>& L, size_t& i, int indent) {
    Node n = Node::map(false);
    while (i < L.size() && L[i].indent == indent && !is_seq_item(L[i].text)) {
        const std::string& t = L[i].text;
        const size_t c = key_sep(t);
        if (c == std::string::npos)
            throw std::runtime_error("yaml: expected 'key: value', line " + std::to_string(L[i].lineno));
        bool q;
        const std::string key = unquote(trim(t.substr(0, c)), &q);
        const std::string rest = trim(t.substr(c + 1));
        const int lineno = L[i].lineno;
        ++i;
        if (!rest.empty()) {
            n[key] = inline_value(rest, lineno);
        } else if (i < L.size() && (L[i].indent > indent ||
                                    (L[i].indent == indent && is_seq_item(L[i].text)))) {
            n[key] = parse_block(L, i, L[i].indent);
        } else {
            n[key] = Node();
        }
    }
    if (i < L.size() && L[i].indent > indent)
        throw std::runtime_error("yaml: unexpected indentation, line " + std::to_string(L[i].lineno));
    return n;
}

Node parse_block(std::vector<Line>& L, size_t& i, int indent) {
    if (is_seq_item(L[i].text)) return parse_seq(L, i, indent);
    return parse_map(L, i, indent);
}

bool needs_quotes(const std::string& s) {
    if (s.empty()) return true;
    if (std::isspace((unsigned char)s.front()) || std::isspace((unsigned char)s.back())) return true;
    if (std::strchr("{}[],&*!|>'\"%@`#-?:", s.front()) && !(s.front() == '-' && s.size() > 1 &&
                                                          (std::isdigit((unsigned char)s[1]) || s[1] == '.')))
        return true;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (c == '\n' || c == '\t' || c == '\r' || c == '"') return true;
        if (c == ':' && (i + 1 == s.size() || s[i + 1] == ' ')) return true;
        if (c == '#' && i > 0 && s[i - 1] == ' ') return true;
    }
    return false;
}

std::string quote(const std::string& s) {
    std::string o = "\"";
    for (char c : s) {
        switch (c) {
            case '\n': o += "\\n"; break;
            case '\t': o += "\\t"; break;
            case '\r': o += "\\r"; break;
            case '"': o += "\\\""; break;
            case '\\': o += "\\\\"; break;
            default: o += c;
        }
    }
    return o + "\"";
}

std::string scalar_text(const Node& n) {
    if (n.is_null()) return "~";
    const std::string& s = n.str();
    return (n.quoted() && needs_quotes(s)) || needs_quotes(s) ? quote(s) : s;
}

std::string flow_text(const Node& n) {
    if (n.is_map()) {
        std::string o = "{";
        bool first = true;
        for (auto& kv : n.items()) {
            if (!first) o += ", ";
            first = false;
            o += (needs_quotes(kv.first) ? quote(kv.first) : kv.first) + ": " + flow_text(kv.second);
        }
        return o + "}";
    }
    if (n.is_seq()) {
        std::string o = "[";
        for (size_t i = 0; i < n.size(); ++i) o += (i ? ", " : "") + flow_text(n.at(i));
        return o + "]";
    }
    return scalar_text(n);
}

void emit(const Node& n, int indent, std::string& o);

void emit_value_after_key(const Node& v, int indent, std::string& o) {
    if (v.is_scalar() || v.is_null()) {
        o += " " + scalar_text(v) + "\n";
    } else if (v.flow() || v.size() == 0) {
        o += " " + flow_text(v) + "\n";
    } else {
        o += "\n";
        emit(v, indent + 2, o);
    }
}

void emit(const Node& n, int indent, std::string& o) {
    const std::string pad(indent, ' ');
    if (n.is_map()) {
        for (auto& kv : n.items()) {
            o += pad + (needs_quotes(kv.first) ? quote(kv.first) : kv.first) + ":";
            emit_value_after_key(kv.second, indent, o);
        }
    } else if (n.is_seq()) {
        for (auto& v : n.elements()) {
            o += pad + "-";
            emit_value_after_key(v, indent, o);
        }
    } else {
        o += pad + scalar_text(n) + "\n";
    }
}

}  // namespace

Node parse(const std::string& text) {
    std::vector<Line> L = split_lines(text);
    if (L.empty()) return Node();
    size_t i = 0;
    Node n = parse_block(L, i, L[0].indent);
    if (i != L.size()) throw std::runtime_error("yaml: unexpected content, line " + std::to_string(L[i].lineno));
    return n;
}

Node load_file(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("yaml: cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

std::string dump(const Node& root) {
    std::string o;
    emit(root, 0, o);
    return o;
}

void save_file(const Node& root, const std::string& path) {
    std::ofstream f(path);
    if (!f) throw std::runtime_error("yaml: cannot write " + path);
    f << dump(root);
}

const Node* find(const Node& root, const std::string& path) {
    const Node* cur = &root;
    size_t a = 0;
    while (a <= path.size()) {
        size_t b = path.find('/', a);
        if (b == std::string::npos) b = path.size();
        const std::string key = path.substr(a, b - a);
        if (!cur->has(key)) return nullptr;
        cur = &cur->at(key);
        a = b + 1;
    }
    return cur;
}

}  // namespace lc2yaml
