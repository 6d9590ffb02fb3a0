// toml.hpp -- a TOML reader for the input files of the MD front end.
//
// Supports what the reference's inputs use: comments, bare / quoted / dotted
// keys (including non-ASCII quoted keys such as "σ"), [tables],
// [[arrays.of.tables]] (nested under the last element of enclosing arrays),
// inline tables, arrays spanning lines, basic and literal strings, integers
// and floats with '_' separators, booleans.  Dates and multi-line strings
// are not supported.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace ocmtoml {

struct Value;
using Array = std::vector<Value>;
using Table = std::vector<std::pair<std::string, Value>>;  // insertion order kept

struct Value {
    enum Kind { None, Bool, Int, Float, String, ArrayK, TableK };
    Kind kind = None;
    bool b = false;
    long long i = 0;
    double f = 0.0;
    std::string s;
    std::shared_ptr<Array> a;
    std::shared_ptr<Table> t;
    int line = 0;

    static Value table();
    static Value array();

    bool is_table() const { return kind == TableK; }
    bool is_array() const { return kind == ArrayK; }
    bool is_string() const { return kind == String; }
    bool is_number() const { return kind == Int || kind == Float; }
    bool contains(const std::string& key) const;
    const Value* find(const std::string& key) const;
    const Value& at(const std::string& key) const;          // throws if missing
    Value* find_mut(const std::string& key);
    Value& insert(const std::string& key, Value v);          // throws on duplicate

    double as_double() const;
    long long as_int() const;
    bool as_bool() const;
    const std::string& as_string() const;
    const Array& as_array() const;
    const Table& as_table() const;
    std::string where() const;
};

struct ParseError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

Value parse(const std::string& text, const std::string& source = "<string>");
Value parse_file(const std::string& path);

}  // namespace ocmtoml
