// A TOML 1.0 reader for the solver's configuration files.
//
// The reference reads its configs with toml++ 3.4.0 (vendored,
// external/tomlplusplus) through the helpers of src/io/toml/
// toml_utilities.cpp.  This reader builds the same document model -- tables
// whose keys iterate in sorted order (toml::table is an ordered map), arrays,
// strings, 64-bit integers, doubles and booleans -- and converts values with
// toml++'s node::value<T>() rules (toml.hpp:5617-5730):
//   as double:  float as is; integer when |i| <= 2^53; nothing else
//   as integer: integer as is; float when finite and integral; bool as 0/1
//   as string:  strings only
// Offset date-times, local date-times, local dates and local times are read
// and validated but convert to nothing (as in toml++, where they are their
// own node types).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

namespace cfdg::toml {

// Syntax error: "line L: description".
struct ParseError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Node;
using NodePtr = std::shared_ptr<Node>;

struct Node {
    enum class Kind { Table, Array, String, Integer, Float, Boolean, DateTime };
    Kind kind = Kind::Table;
    std::map<std::string, NodePtr> table;  // sorted keys, as toml::table iterates
    std::vector<NodePtr> array;
    std::string str;  // a string's value, or a date / time as written
    std::int64_t i = 0;
    double f = 0.0;
    bool b = false;
    // how a table came to be: by a [header], implicitly (as the parent of a
    // header or dotted key), inline ({...}), or as an element of [[array]]
    bool explicit_header = false, inline_table = false, dotted = false;
    bool table_array = false;  // an array created by [[header]]

    bool is_table() const { return kind == Kind::Table; }
    bool is_array() const { return kind == Kind::Array; }
    const Node* get(const std::string& key) const;  // table member or null
    bool contains(const std::string& key) const { return get(key) != nullptr; }

    std::optional<double> as_double() const;
    std::optional<std::int64_t> as_integer() const;
    std::optional<std::string> as_string() const;
};

// Parses a whole document; throws ParseError.
NodePtr parse(const std::string& text);

}  // namespace cfdg::toml
