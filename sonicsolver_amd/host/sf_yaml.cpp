// sf_yaml.cpp -- parser of the YAML subset described in sf_yaml.h.
#include "sf_yaml.h"

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace sfamd::host::yaml {

namespace {

[[noreturn]] void fail(const std::string& source, int line, const std::string& what) {
    throw std::runtime_error(source + ":" + std::to_string(line) + ": " + what);
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// Remove a '#' comment that is outside quotes and starts a token.
std::string strip_comment(const std::string& s) {
    char q = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (q) {
            if (c == '\\' && q == '"') {
                ++i;
                continue;
            }
            if (c == q) q = 0;
            continue;
        }
        if (c == '"' || c == '\'') q = c;
        else if (c == '#' && (i == 0 || std::isspace((unsigned char)s[i - 1]))) return s.substr(0, i);
    }
    return s;
}

// Flow-style parser over one string (a value that may span lines).
class Flow {
public:
    Flow(const std::string& s, const std::string& source, int line)
        : s_(s), src_(source), line_(line) {}

    Node value(bool in_flow) {
        skip();
        if (p_ >= s_.size()) return Node{};
        const char c = s_[p_];
        if (c == '{') return map();
        if (c == '[') return seq();
        if (c == '"' || c == '\'') return quoted();
        return plain(in_flow);
    }

    void expect_end() {
        skip();
        if (p_ != s_.size()) fail(src_, line_, "unexpected text after value: '" + s_.substr(p_) + "'");
    }

private:
    void skip() {
        while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) {
            if (s_[p_] == '\n') ++line_;
            ++p_;
        }
    }

    Node map() {
        Node n;
        n.kind = Node::Map;
        n.line = line_;
        ++p_;
        skip();
        if (p_ < s_.size() && s_[p_] == '}') {
            ++p_;
            return n;
        }
        for (;;) {
            skip();
            Node k = (p_ < s_.size() && (s_[p_] == '"' || s_[p_] == '\'')) ? quoted() : key();
            skip();
            if (p_ >= s_.size() || s_[p_] != ':') fail(src_, line_, "expected ':' in a flow mapping");
            ++p_;
            Node v = value(true);
            n.members.emplace_back(k.text, std::move(v));
            skip();
            if (p_ >= s_.size()) fail(src_, line_, "unterminated flow mapping");
            if (s_[p_] == ',') {
                ++p_;
                continue;
            }
            if (s_[p_] == '}') {
                ++p_;
                return n;
            }
            fail(src_, line_, "expected ',' or '}' in a flow mapping");
        }
    }

    Node seq() {
        Node n;
        n.kind = Node::Seq;
        n.line = line_;
        ++p_;
        skip();
        if (p_ < s_.size() && s_[p_] == ']') {
            ++p_;
            return n;
        }
        for (;;) {
            n.items.push_back(value(true));
            skip();
            if (p_ >= s_.size()) fail(src_, line_, "unterminated flow sequence");
            if (s_[p_] == ',') {
                ++p_;
                continue;
            }
            if (s_[p_] == ']') {
                ++p_;
                return n;
            }
            fail(src_, line_, "expected ',' or ']' in a flow sequence");
        }
    }

    Node quoted() {
        Node n;
        n.kind = Node::Scalar;
        n.quoted = true;
        n.line = line_;
        const char q = s_[p_++];
        std::string out;
        for (;;) {
            if (p_ >= s_.size()) fail(src_, line_, "unterminated quoted string");
            const char c = s_[p_++];
            if (q == '\'' && c == '\'') {
                if (p_ < s_.size() && s_[p_] == '\'') {
                    out += '\'';
                    ++p_;
                    continue;
                }
                break;
            }
            if (q == '"' && c == '"') break;
            if (q == '"' && c == '\\') {
                if (p_ >= s_.size()) fail(src_, line_, "bad escape");
                const char e = s_[p_++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case '"': out += '"'; break;
                    case '\\': out += '\\'; break;
                    case '/': out += '/'; break;
                    default: fail(src_, line_, std::string("unsupported escape \\") + e);
                }
                continue;
            }
            out += c;
        }
        n.text = out;
        return n;
    }

    Node key() {
        Node n;
        n.kind = Node::Scalar;
        n.line = line_;
        const size_t a = p_;
        while (p_ < s_.size() && s_[p_] != ':' && s_[p_] != ',' && s_[p_] != '}') ++p_;
        n.text = trim(s_.substr(a, p_ - a));
        return n;
    }

    Node plain(bool in_flow) {
        Node n;
        n.kind = Node::Scalar;
        n.line = line_;
        const size_t a = p_;
        if (in_flow) {
            while (p_ < s_.size() && s_[p_] != ',' && s_[p_] != ']' && s_[p_] != '}') ++p_;
        } else {
            p_ = s_.size();
        }
        n.text = trim(s_.substr(a, p_ - a));
        if (n.text == "null" || n.text == "~" || n.text.empty()) n.kind = Node::Null;
        return n;
    }

    const std::string& s_;
    std::string src_;
    int line_;
    size_t p_ = 0;
};

struct Line {
    int indent;
    std::string text;
    int number;
};

int bracket_balance(const std::string& s) {
    int depth = 0;
    char q = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (q) {
            if (c == '\\' && q == '"') {
                ++i;
                continue;
            }
            if (c == q) q = 0;
            continue;
        }
        if (c == '"' || c == '\'') q = c;
        else if (c == '[' || c == '{') ++depth;
        else if (c == ']' || c == '}') --depth;
    }
    return depth;
}

// Position of the ':' that ends a block mapping key, or npos.
size_t key_colon(const std::string& s) {
    char q = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (q) {
            if (c == '\\' && q == '"') {
                ++i;
                continue;
            }
            if (c == q) q = 0;
            continue;
        }
        if ((c == '"' || c == '\'') && i == 0) q = c;
        else if (c == '[' || c == '{') return std::string::npos;
        else if (c == ':' && (i + 1 == s.size() || s[i + 1] == ' ' || s[i + 1] == '\t')) return i;
    }
    return std::string::npos;
}

bool is_seq_item(const std::string& t) { return t == "-" || (t.size() > 1 && t[0] == '-' && t[1] == ' '); }

class Block {
public:
    Block(std::vector<Line> lines, const std::string& source)
        : l_(std::move(lines)), src_(source) {}

    Node document() {
        if (l_.empty()) return Node{};
        Node n = block(l_[0].indent);
        if (p_ != l_.size()) fail(src_, l_[p_].number, "bad indentation");
        return n;
    }

private:
    Node block(int indent) {
        if (is_seq_item(l_[p_].text)) return seq(indent);
        if (key_colon(l_[p_].text) != std::string::npos) return map(indent);
        // a lone scalar or flow value
        return inline_value(l_[p_].text, indent);
    }

    // A value on the line at p_ (already stripped of its key); a flow
    // collection continues over the following lines until balanced.
    Node inline_value(std::string text, int indent) {
        const int line = l_[p_].number;
        ++p_;
        while (bracket_balance(text) > 0) {
            if (p_ >= l_.size()) fail(src_, line, "unterminated flow collection");
            text += "\n" + l_[p_].text;
            ++p_;
        }
        (void)indent;
        Flow f(text, src_, line);
        Node n = f.value(false);
        f.expect_end();
        if (n.line == 0) n.line = line;
        return n;
    }

    Node map(int indent) {
        Node n;
        n.kind = Node::Map;
        n.line = l_[p_].number;
        while (p_ < l_.size() && l_[p_].indent == indent && !is_seq_item(l_[p_].text)) {
            const std::string& t = l_[p_].text;
            const size_t c = key_colon(t);
            if (c == std::string::npos) fail(src_, l_[p_].number, "expected 'key: value'");
            std::string key = trim(t.substr(0, c));
            if (key.size() >= 2 && (key[0] == '"' || key[0] == '\'') && key.back() == key[0])
                key = key.substr(1, key.size() - 2);
            for (const auto& m : n.members)
                if (m.first == key) fail(src_, l_[p_].number, "duplicate key '" + key + "'");
            const std::string rest = trim(t.substr(c + 1));
            const int line = l_[p_].number;
            Node v;
            if (!rest.empty()) {
                v = inline_value(rest, indent);
            } else {
                ++p_;
                if (p_ < l_.size() && l_[p_].indent > indent) v = block(l_[p_].indent);
                else if (p_ < l_.size() && l_[p_].indent == indent && is_seq_item(l_[p_].text))
                    v = seq(indent);
                else v.line = line;
            }
            n.members.emplace_back(key, std::move(v));
        }
        if (p_ < l_.size() && l_[p_].indent > indent) fail(src_, l_[p_].number, "bad indentation");
        return n;
    }

    Node seq(int indent) {
        Node n;
        n.kind = Node::Seq;
        n.line = l_[p_].number;
        while (p_ < l_.size() && l_[p_].indent == indent && is_seq_item(l_[p_].text)) {
            std::string rest = l_[p_].text == "-" ? std::string() : trim(l_[p_].text.substr(2));
            if (rest.empty()) {
                ++p_;
                if (p_ < l_.size() && l_[p_].indent > indent) n.items.push_back(block(l_[p_].indent));
                else n.items.emplace_back();
                continue;
            }
            if (key_colon(rest) != std::string::npos || is_seq_item(rest)) {
                // "- key: value" opens a mapping (or "- - x" a sequence)
                // whose first entry sits after the dash
                const std::string& t = l_[p_].text;
                const int inner = indent + (int)t.find_first_not_of(' ', 1);
                l_[p_].indent = inner;
                l_[p_].text = rest;
                n.items.push_back(block(inner));
                continue;
            }
            n.items.push_back(inline_value(rest, indent));
        }
        return n;
    }

    std::vector<Line> l_;
    std::string src_;
    size_t p_ = 0;
};

}  // namespace

const Node* Node::find(const std::string& key) const {
    if (kind != Map) return nullptr;
    for (const auto& m : members)
        if (m.first == key) return &m.second;
    return nullptr;
}

bool Node::isNumber() const {
    if (kind != Scalar || quoted || text.empty()) return false;
    const char* b = text.c_str();
    char* e = nullptr;
    errno = 0;
    std::strtod(b, &e);
    return e && *e == '\0' && errno == 0;
}

double Node::number() const {
    if (!isNumber()) throw std::runtime_error("line " + std::to_string(line) + ": '" + text +
                                              "' is not a number");
    return std::strtod(text.c_str(), nullptr);
}

long long Node::integer() const {
    const double v = number();
    if (v != std::floor(v)) throw std::runtime_error("line " + std::to_string(line) + ": '" + text +
                                                     "' is not an integer");
    return (long long)v;
}

bool Node::boolean() const {
    if (kind == Scalar && !quoted) {
        if (text == "true" || text == "True" || text == "TRUE" || text == "yes") return true;
        if (text == "false" || text == "False" || text == "FALSE" || text == "no") return false;
    }
    throw std::runtime_error("line " + std::to_string(line) + ": '" + text + "' is not a boolean");
}

const std::string& Node::string() const {
    if (kind != Scalar) throw std::runtime_error("line " + std::to_string(line) +
                                                 ": expected a scalar");
    return text;
}

Node parse(const std::string& text, const std::string& source) {
    std::vector<Line> lines;
    std::istringstream in(text);
    std::string raw;
    int number = 0;
    while (std::getline(in, raw)) {
        ++number;
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        if (raw.find('\t') != std::string::npos && raw.find_first_not_of(" \t") != std::string::npos &&
            raw.find_first_not_of(' ') < raw.size() && raw[raw.find_first_not_of(' ')] == '\t')
            fail(source, number, "tab indentation");
        const std::string body = strip_comment(raw);
        const std::string t = trim(body);
        if (t.empty() || t == "---") continue;
        const int indent = (int)body.find_first_not_of(' ');
        lines.push_back({indent, t, number});
    }
    Block b(std::move(lines), source);
    return b.document();
}

Node parseFile(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error(path + ": cannot open");
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str(), path);
}

}  // namespace sfamd::host::yaml
