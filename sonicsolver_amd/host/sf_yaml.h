// sf_yaml.h -- the YAML subset SonicFile documents use: block mappings and
// sequences by indentation, flow collections ([..], {..}, possibly over
// several lines, so a JSON document parses too), plain / single / double
// quoted scalars and '#' comments.  No anchors, tags or multi-documents.
// Map members keep file order.
#pragma once

#include <string>
#include <utility>
#include <vector>

namespace sfamd::host::yaml {

struct Node {
    enum Kind { Null, Scalar, Seq, Map };
    Kind kind = Null;
    std::string text;      // scalar text, unquoted
    bool quoted = false;
    int line = 0;          // 1-based line of the node in its file
    std::vector<Node> items;
    std::vector<std::pair<std::string, Node>> members;

    bool isNull() const { return kind == Null; }
    bool isMap() const { return kind == Map; }
    bool isSeq() const { return kind == Seq; }
    bool isScalar() const { return kind == Scalar; }
    const Node* find(const std::string& key) const;
    // number / bool / string of a scalar; throw std::runtime_error otherwise
    bool isNumber() const;
    double number() const;
    long long integer() const;
    bool boolean() const;
    const std::string& string() const;
};

// Parse a document; errors throw std::runtime_error naming `source` and the
// line.
Node parse(const std::string& text, const std::string& source);
Node parseFile(const std::string& path);

}  // namespace sfamd::host::yaml
