// toml.cpp -- recursive-descent reader for the TOML subset in toml.hpp.
#include "toml.hpp"

#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>

namespace ocmtoml {

Value Value::table() {
    Value v;
    v.kind = TableK;
    v.t = std::make_shared<Table>();
    return v;
}
Value Value::array() {
    Value v;
    v.kind = ArrayK;
    v.a = std::make_shared<Array>();
    return v;
}

std::string Value::where() const { return line > 0 ? " (line " + std::to_string(line) + ")" : ""; }

bool Value::contains(const std::string& key) const { return find(key) != nullptr; }

const Value* Value::find(const std::string& key) const {
    if (kind != TableK) return nullptr;
    for (const auto& kv : *t)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

Value* Value::find_mut(const std::string& key) {
    if (kind != TableK) return nullptr;
    for (auto& kv : *t)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

const Value& Value::at(const std::string& key) const {
    const Value* v = find(key);
    if (!v) throw std::out_of_range("[error] key \"" + key + "\" not found" + where());
    return *v;
}

Value& Value::insert(const std::string& key, Value v) {
    if (kind != TableK) throw ParseError("[error] inserting into a non-table value" + where());
    if (find(key)) throw ParseError("[error] duplicate key \"" + key + "\"" + v.where());
    t->emplace_back(key, std::move(v));
    return t->back().second;
}

double Value::as_double() const {
    if (kind == Float) return f;
    if (kind == Int) return static_cast<double>(i);
    throw std::runtime_error("[error] expected a number" + where());
}
long long Value::as_int() const {
    if (kind == Int) return i;
    throw std::runtime_error("[error] expected an integer" + where());
}
bool Value::as_bool() const {
    if (kind == Bool) return b;
    throw std::runtime_error("[error] expected a boolean" + where());
}
const std::string& Value::as_string() const {
    if (kind == String) return s;
    throw std::runtime_error("[error] expected a string" + where());
}
const Array& Value::as_array() const {
    if (kind == ArrayK) return *a;
    throw std::runtime_error("[error] expected an array" + where());
}
const Table& Value::as_table() const {
    if (kind == TableK) return *t;
    throw std::runtime_error("[error] expected a table" + where());
}

namespace {

class Parser {
public:
    Parser(const std::string& text, const std::string& src) : s_(text), src_(src) {}

    Value run() {
        Value root = Value::table();
        Value* cur = &root;
        while (true) {
            skip_ws_nl_comments();
            if (eof()) break;
            if (peek() == '[') {
                const bool aot = s_.compare(p_, 2, "[[") == 0;
                p_ += aot ? 2 : 1;
                std::vector<std::string> path = parse_key_path();
                skip_ws();
                expect(']');
                if (aot) expect(']');
                cur = aot ? open_array_table(root, path) : open_table(root, path);
                end_of_line();
            } else {
                parse_keyval(*cur);
                end_of_line();
            }
        }
        return root;
    }

private:
    const std::string& s_;
    std::string src_;
    size_t p_ = 0;
    int line_ = 1;

    [[noreturn]] void fail(const std::string& m) const {
        throw ParseError("[error] " + src_ + ":" + std::to_string(line_) + ": " + m);
    }
    bool eof() const { return p_ >= s_.size(); }
    char peek() const { return eof() ? '\0' : s_[p_]; }
    char get() {
        if (eof()) fail("unexpected end of input");
        const char c = s_[p_++];
        if (c == '\n') ++line_;
        return c;
    }
    void expect(char c) {
        if (peek() != c) fail(std::string("expected '") + c + "'");
        get();
    }
    void skip_ws() {
        while (!eof() && (peek() == ' ' || peek() == '\t' || peek() == '\r')) ++p_;
    }
    void skip_comment() {
        if (peek() == '#')
            while (!eof() && peek() != '\n') ++p_;
    }
    void skip_ws_nl_comments() {
        while (!eof()) {
            skip_ws();
            skip_comment();
            if (peek() == '\n') get();
            else break;
        }
    }
    void end_of_line() {
        skip_ws();
        skip_comment();
        if (!eof()) {
            if (peek() != '\n') fail("unexpected trailing characters");
            get();
        }
    }

    std::string parse_simple_key() {
        skip_ws();
        if (peek() == '"') return parse_basic_string();
        if (peek() == '\'') return parse_literal_string();
        std::string k;
        while (!eof()) {
            const char c = peek();
            if (std::isalnum(static_cast<unsigned char>(c)) || c == '_' || c == '-') { k += c; ++p_; }
            else break;
        }
        if (k.empty()) fail("expected a key");
        return k;
    }
    std::vector<std::string> parse_key_path() {
        std::vector<std::string> path{parse_simple_key()};
        while (true) {
            skip_ws();
            if (peek() != '.') break;
            ++p_;
            path.push_back(parse_simple_key());
        }
        return path;
    }

    // descend through path[0..n-2], entering the last element of arrays of tables
    Value* descend(Value& root, const std::vector<std::string>& path, size_t upto) {
        Value* cur = &root;
        for (size_t k = 0; k < upto; ++k) {
            Value* nxt = cur->find_mut(path[k]);
            if (!nxt) nxt = &cur->insert(path[k], Value::table());
            if (nxt->is_array()) {
                if (nxt->a->empty() || !nxt->a->back().is_table()) fail("key \"" + path[k] + "\" is not a table");
                nxt = &nxt->a->back();
            } else if (!nxt->is_table()) {
                fail("key \"" + path[k] + "\" is not a table");
            }
            cur = nxt;
        }
        return cur;
    }
    Value* open_table(Value& root, const std::vector<std::string>& path) {
        Value* parent = descend(root, path, path.size() - 1);
        Value* t = parent->find_mut(path.back());
        if (!t) {
            Value nt = Value::table();
            nt.line = line_;
            t = &parent->insert(path.back(), nt);
        }
        if (!t->is_table()) fail("key \"" + path.back() + "\" is not a table");
        return t;
    }
    Value* open_array_table(Value& root, const std::vector<std::string>& path) {
        Value* parent = descend(root, path, path.size() - 1);
        Value* arr = parent->find_mut(path.back());
        if (!arr) arr = &parent->insert(path.back(), Value::array());
        if (!arr->is_array()) fail("key \"" + path.back() + "\" is not an array of tables");
        Value nt = Value::table();
        nt.line = line_;
        arr->a->push_back(nt);
        return &arr->a->back();
    }

    void parse_keyval(Value& table) {
        const int l = line_;
        std::vector<std::string> path = parse_key_path();
        skip_ws();
        expect('=');
        skip_ws();
        Value v = parse_value();
        v.line = l;
        Value* parent = descend(table, path, path.size() - 1);
        if (parent->find(path.back())) fail("duplicate key \"" + path.back() + "\"");
        parent->insert(path.back(), std::move(v));
    }

    static void append_utf8(std::string& out, unsigned cp) {
        if (cp < 0x80) out += static_cast<char>(cp);
        else if (cp < 0x800) { out += static_cast<char>(0xC0 | (cp >> 6)); out += static_cast<char>(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) {
            out += static_cast<char>(0xE0 | (cp >> 12));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        } else {
            out += static_cast<char>(0xF0 | (cp >> 18));
            out += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        }
    }
    std::string parse_basic_string() {
        expect('"');
        std::string out;
        while (true) {
            if (eof()) fail("unterminated string");
            char c = get();
            if (c == '"') break;
            if (c == '\n') fail("newline in string");
            if (c == '\\') {
                const char e = get();
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case 'u': case 'U': {
                    const int nd = e == 'u' ? 4 : 8;
                    unsigned cp = 0;
                    for (int k = 0; k < nd; ++k) {
                        const char h = get();
                        if (!std::isxdigit(static_cast<unsigned char>(h))) fail("bad unicode escape");
                        cp = cp * 16 + static_cast<unsigned>(std::isdigit(static_cast<unsigned char>(h)) ? h - '0' : (std::tolower(h) - 'a' + 10));
                    }
                    append_utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
                }
            } else {
                out += c;
            }
        }
        return out;
    }
    std::string parse_literal_string() {
        expect('\'');
        std::string out;
        while (true) {
            if (eof()) fail("unterminated string");
            const char c = get();
            if (c == '\'') break;
            if (c == '\n') fail("newline in string");
            out += c;
        }
        return out;
    }

    void skip_array_ws() {
        while (!eof()) {
            skip_ws();
            skip_comment();
            if (peek() == '\n') get();
            else break;
        }
    }

    Value parse_value() {
        const char c = peek();
        Value v;
        v.line = line_;
        if (c == '"') { v.kind = Value::String; v.s = parse_basic_string(); return v; }
        if (c == '\'') { v.kind = Value::String; v.s = parse_literal_string(); return v; }
        if (c == '[') {
            get();
            v = Value::array();
            v.line = line_;
            while (true) {
                skip_array_ws();
                if (peek() == ']') { get(); break; }
                v.a->push_back(parse_value());
                skip_array_ws();
                if (peek() == ',') { get(); continue; }
                if (peek() == ']') { get(); break; }
                fail("expected ',' or ']' in array");
            }
            return v;
        }
        if (c == '{') {
            get();
            v = Value::table();
            v.line = line_;
            skip_ws();
            if (peek() == '}') { get(); return v; }
            while (true) {
                parse_keyval(v);
                skip_ws();
                if (peek() == ',') { get(); skip_ws(); continue; }
                if (peek() == '}') { get(); break; }
                fail("expected ',' or '}' in inline table");
            }
            return v;
        }
        if (s_.compare(p_, 4, "true") == 0) { p_ += 4; v.kind = Value::Bool; v.b = true; return v; }
        if (s_.compare(p_, 5, "false") == 0) { p_ += 5; v.kind = Value::Bool; v.b = false; return v; }
        // number
        std::string tok;
        while (!eof()) {
            const char d = peek();
            if (std::isalnum(static_cast<unsigned char>(d)) || d == '+' || d == '-' || d == '.' || d == '_') { tok += d; ++p_; }
            else break;
        }
        if (tok.empty()) fail("expected a value");
        std::string clean;
        for (char d : tok)
            if (d != '_') clean += d;
        const bool is_float = clean.find_first_of(".eE") != std::string::npos ||
                              clean.find("inf") != std::string::npos || clean.find("nan") != std::string::npos;
        try {
            size_t used = 0;
            if (is_float) {
                v.kind = Value::Float;
                if (clean == "inf" || clean == "+inf") v.f = HUGE_VAL;
                else if (clean == "-inf") v.f = -HUGE_VAL;
                else if (clean == "nan" || clean == "+nan" || clean == "-nan") v.f = std::nan("");
                else { v.f = std::stod(clean, &used); if (used != clean.size()) fail("bad number \"" + tok + "\""); }
            } else {
                v.kind = Value::Int;
                if (clean.size() > 2 && clean[0] == '0' && (clean[1] == 'x' || clean[1] == 'o' || clean[1] == 'b')) {
                    const int base = clean[1] == 'x' ? 16 : (clean[1] == 'o' ? 8 : 2);
                    v.i = std::stoll(clean.substr(2), &used, base);
                    used += 2;
                } else {
                    v.i = std::stoll(clean, &used, 10);
                }
                if (used != clean.size()) fail("bad number \"" + tok + "\"");
            }
        } catch (const ParseError&) {
            throw;
        } catch (const std::exception&) {
            fail("bad number \"" + tok + "\"");
        }
        return v;
    }
};

}  // namespace

Value parse(const std::string& text, const std::string& source) { return Parser(text, source).run(); }

Value parse_file(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("[error] cannot open file " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return parse(ss.str(), path);
}

}  // namespace ocmtoml
