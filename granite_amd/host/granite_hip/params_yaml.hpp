// The subset of YAML that GRANITE's parameter files use (benchmarks/*/params.yaml): nested
// block mappings by indentation, scalars, flow sequences [a, b, c], block sequences of
// mappings ("- key: value" items), folded (>) and literal (|) block scalars with their
// chomping indicators (the descriptions of benchmarks/B5_star/params.yaml:10 and
// benchmarks/validation_tests.yaml), and # comments. It stands in for yaml-cpp, which the
// reference driver links (src/main.cpp:730-830) and which is not available here.
//
// Header-only C++17. Errors throw std::runtime_error with the line number.
#pragma once

#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace granite_hip {

struct YNode {
    enum Kind { Null, Scalar, Map, Seq } kind = Null;
    std::string text;
    std::vector<std::pair<std::string, YNode>> map;
    std::vector<YNode> seq;

    explicit operator bool() const { return kind != Null; }
    const YNode& operator[](const std::string& key) const {
        static const YNode none;
        if (kind == Map)
            for (const auto& kv : map)
                if (kv.first == key) return kv.second;
        return none;
    }
    const YNode& operator[](size_t i) const {
        static const YNode none;
        return (kind == Seq && i < seq.size()) ? seq[i] : none;
    }
    size_t size() const { return kind == Seq ? seq.size() : (kind == Map ? map.size() : 0); }
    double as_double() const {
        char* end = nullptr;
        const double v = std::strtod(text.c_str(), &end);
        if (kind != Scalar || end == text.c_str() || *end) throw std::runtime_error("yaml: not a number: '" + text + "'");
        return v;
    }
    int as_int() const {
        const double v = as_double();
        if (v != (double)(int)v) throw std::runtime_error("yaml: not an integer: '" + text + "'");
        return (int)v;
    }
    const std::string& as_string() const {
        if (kind != Scalar) throw std::runtime_error("yaml: not a scalar");
        return text;
    }
};

namespace yaml_detail {

struct Line {
    int indent;
    std::string text;
    int number;
    bool block = false;     // "key: >" or "key: |": the value is block_text
    std::string block_text;
};

inline std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

inline std::string strip_comment(const std::string& s) {
    bool sq = false, dq = false;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (c == '\'' && !dq) sq = !sq;
        else if (c == '"' && !sq) dq = !dq;
        else if (c == '#' && !sq && !dq && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) return s.substr(0, i);
    }
    return s;
}

inline bool is_item_text(const std::string& t) { return t == "-" || t.rfind("- ", 0) == 0; }

inline YNode scalar(std::string t) {
    t = trim(t);
    if (t.size() >= 2 && ((t.front() == '"' && t.back() == '"') || (t.front() == '\'' && t.back() == '\'')))
        t = t.substr(1, t.size() - 2);
    YNode n;
    n.kind = YNode::Scalar;
    n.text = t;
    return n;
}

inline YNode value(const std::string& raw, int line) {
    const std::string t = trim(raw);
    if (t.empty()) return YNode();
    if (t.front() == '[') {
        if (t.back() != ']') throw std::runtime_error("yaml line " + std::to_string(line) + ": unterminated [");
        YNode n;
        n.kind = YNode::Seq;
        const std::string body = t.substr(1, t.size() - 2);
        std::string item;
        std::stringstream ss(body);
        while (std::getline(ss, item, ','))
            if (!trim(item).empty()) n.seq.push_back(scalar(item));
        return n;
    }
    return scalar(t);
}

class Parser {
public:
    explicit Parser(std::vector<Line> lines) : L_(std::move(lines)) {}
    YNode parse() {
        size_t i = 0;
        if (L_.empty()) return YNode();
        YNode n = block(i, L_[0].indent);
        if (i != L_.size()) throw std::runtime_error("yaml line " + std::to_string(L_[i].number) + ": bad indentation");
        return n;
    }

private:
    std::vector<Line> L_;

    static bool is_item(const Line& l) { return l.text == "-" || l.text.rfind("- ", 0) == 0; }

    YNode block(size_t& i, int indent) {
        if (is_item(L_[i])) return sequence(i, indent);
        return mapping(i, indent);
    }

    YNode sequence(size_t& i, int indent) {
        YNode n;
        n.kind = YNode::Seq;
        while (i < L_.size() && L_[i].indent == indent && is_item(L_[i])) {
            const std::string rest = L_[i].text.size() > 1 ? L_[i].text.substr(2) : std::string();
            const int sub = indent + 2;
            if (trim(rest).find(':') != std::string::npos && trim(rest).front() != '[') {
                // "- key: value": the item is a mapping whose first line is `rest`
                L_[i].indent = sub;
                L_[i].text = trim(rest);
                n.seq.push_back(mapping(i, sub));
            } else if (!trim(rest).empty()) {
                n.seq.push_back(value(rest, L_[i].number));
                ++i;
            } else {
                ++i;
                n.seq.push_back(i < L_.size() && L_[i].indent > indent ? block(i, L_[i].indent) : YNode());
            }
        }
        return n;
    }

    YNode mapping(size_t& i, int indent) {
        YNode n;
        n.kind = YNode::Map;
        while (i < L_.size() && L_[i].indent == indent && !is_item(L_[i])) {
            const Line& l = L_[i];
            const size_t c = l.text.find(':');
            if (c == std::string::npos) throw std::runtime_error("yaml line " + std::to_string(l.number) + ": expected 'key: value'");
            const std::string key = trim(l.text.substr(0, c));
            const std::string rest = l.text.substr(c + 1);
            ++i;
            if (l.block) {
                YNode v;
                v.kind = YNode::Scalar;
                v.text = l.block_text;
                n.map.emplace_back(key, v);
            } else if (!trim(rest).empty()) {
                n.map.emplace_back(key, value(rest, l.number));
            } else if (i < L_.size() && (L_[i].indent > indent || (L_[i].indent == indent && is_item(L_[i])))) {
                n.map.emplace_back(key, block(i, L_[i].indent));
            } else {
                n.map.emplace_back(key, YNode());
            }
        }
        return n;
    }
};

} // namespace yaml_detail

namespace yaml_detail {

// "key: >", "key: |-" ...: the block-scalar header of a mapping line, or '\0'
inline char block_header(const std::string& text, char* chomp) {
    const size_t c = text.find(':');
    if (c == std::string::npos) return '\0';
    const std::string v = trim(text.substr(c + 1));
    if (v.empty() || (v[0] != '>' && v[0] != '|')) return '\0';
    *chomp = ' ';
    for (size_t q = 1; q < v.size(); ++q) {
        if (v[q] == '-' || v[q] == '+') *chomp = v[q];
        else if (!(v[q] >= '1' && v[q] <= '9')) return '\0';
    }
    return v[0];
}

// The body of a block scalar: the lines after the header that are blank or indented deeper
// than the key, with the indentation of its first line removed. Literal (|) keeps the line
// breaks, folded (>) joins lines with spaces (each blank line becomes one line break); chomping
// '-' strips the final line break, '+' keeps trailing blank lines, clip keeps one.
inline std::string block_body(const std::vector<std::string>& raw, size_t& r, int key_indent, char style, char chomp) {
    std::vector<std::string> body;
    int ind = -1;
    while (r < raw.size()) {
        const std::string& l = raw[r];
        const bool blank = trim(l).empty();
        int li = 0;
        while (li < (int)l.size() && l[li] == ' ') ++li;
        if (!blank && li <= key_indent) break;
        if (!blank && ind < 0) ind = li;
        body.push_back(blank ? std::string() : l.substr(std::min<size_t>((size_t)std::max(ind, 0), l.size())));
        ++r;
    }
    size_t trailing = 0;
    while (!body.empty() && body.back().empty()) {
        body.pop_back();
        ++trailing;
    }
    std::string out;
    for (size_t q = 0; q < body.size(); ++q) {
        if (style == '|') {
            if (q > 0) out += "\n";
        } else if (body[q].empty()) {
            out += "\n"; // a blank line of a folded scalar is one line break
        } else if (q > 0 && !body[q - 1].empty()) {
            out += " ";
        }
        out += body[q];
    }
    if (chomp == '+') out += std::string(trailing + (body.empty() ? 0 : 1), '\n');
    else if (chomp == ' ' && !body.empty()) out += "\n";
    return out;
}

} // namespace yaml_detail

inline YNode parse_yaml(const std::string& text) {
    std::vector<yaml_detail::Line> lines;
    std::vector<std::string> raw;
    {
        std::stringstream ss(text);
        std::string l;
        while (std::getline(ss, l)) {
            if (!l.empty() && l.back() == '\r') l.pop_back();
            raw.push_back(l);
        }
    }
    size_t r = 0;
    while (r < raw.size()) {
        const int number = (int)r + 1;
        const std::string s = yaml_detail::strip_comment(raw[r++]);
        if (yaml_detail::trim(s).empty()) continue;
        int ind = 0;
        while (ind < (int)s.size() && s[ind] == ' ') ++ind;
        if (ind < (int)s.size() && s[ind] == '\t') throw std::runtime_error("yaml line " + std::to_string(number) + ": tab indentation");
        yaml_detail::Line line{ind, yaml_detail::trim(s), number, false, std::string()};
        char chomp = ' ';
        const char style = yaml_detail::is_item_text(line.text) ? '\0' : yaml_detail::block_header(line.text, &chomp);
        if (style) {
            line.block = true;
            line.block_text = yaml_detail::block_body(raw, r, ind, style, chomp);
        }
        lines.push_back(std::move(line));
    }
    return yaml_detail::Parser(std::move(lines)).parse();
}

inline YNode load_yaml_file(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse_yaml(ss.str());
}

} // namespace granite_hip
