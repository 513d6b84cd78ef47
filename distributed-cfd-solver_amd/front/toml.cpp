// TOML 1.0 reader (document model and value conversions in toml.hpp).
#include "toml.hpp"

#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <limits>

namespace cfdg::toml {

const Node* Node::get(const std::string& key) const {
    if (kind != Kind::Table) return nullptr;
    const auto it = table.find(key);
    return it == table.end() ? nullptr : it->second.get();
}

std::optional<double> Node::as_double() const {
    if (kind == Kind::Float) return f;
    if (kind == Kind::Integer) {
        constexpr std::int64_t kWhole = std::int64_t{1} << std::numeric_limits<double>::digits;
        if (i < -kWhole || i > kWhole) return std::nullopt;
        return static_cast<double>(i);
    }
    return std::nullopt;
}

std::optional<std::int64_t> Node::as_integer() const {
    if (kind == Kind::Integer) return i;
    if (kind == Kind::Boolean) return b ? 1 : 0;
    if (kind == Kind::Float) {
        if (std::isfinite(f) && static_cast<double>(static_cast<std::int64_t>(f)) == f)
            return static_cast<std::int64_t>(f);
        return std::nullopt;
    }
    return std::nullopt;
}

std::optional<std::string> Node::as_string() const {
    if (kind == Kind::String) return str;
    return std::nullopt;
}

namespace {

NodePtr make(Node::Kind k) {
    auto n = std::make_shared<Node>();
    n->kind = k;
    return n;
}

bool is_bare(char c) {
    return (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || (c >= '0' && c <= '9') || c == '_' || c == '-';
}
bool is_digit(char c) { return c >= '0' && c <= '9'; }

void append_utf8(std::string& s, std::uint32_t cp) {
    if (cp < 0x80) {
        s += static_cast<char>(cp);
    } else if (cp < 0x800) {
        s += static_cast<char>(0xC0 | (cp >> 6));
        s += static_cast<char>(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
        s += static_cast<char>(0xE0 | (cp >> 12));
        s += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
        s += static_cast<char>(0x80 | (cp & 0x3F));
    } else {
        s += static_cast<char>(0xF0 | (cp >> 18));
        s += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
        s += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
        s += static_cast<char>(0x80 | (cp & 0x3F));
    }
}

class Parser {
public:
    explicit Parser(const std::string& t) : s_(t) {}

    NodePtr run() {
        root_ = make(Node::Kind::Table);
        Node* cur = root_.get();
        while (true) {
            skip_ws();
            if (eof()) break;
            const char c = peek();
            if (c == '#') {
                skip_comment();
            } else if (c == '\n' || c == '\r') {
                newline();
                continue;
            } else if (c == '[') {
                cur = header();
            } else {
                keyval(cur);
            }
            end_of_line();
        }
        return root_;
    }

private:
    const std::string& s_;
    std::size_t p_ = 0;
    int line_ = 1;
    NodePtr root_;

    [[noreturn]] void fail(const std::string& what) const {
        throw ParseError("line " + std::to_string(line_) + ": " + what);
    }
    bool eof() const { return p_ >= s_.size(); }
    char peek(std::size_t k = 0) const { return p_ + k < s_.size() ? s_[p_ + k] : '\0'; }
    void skip_ws() {
        while (!eof() && (peek() == ' ' || peek() == '\t')) ++p_;
    }
    void skip_comment() {
        while (!eof() && peek() != '\n') {
            const unsigned char c = static_cast<unsigned char>(peek());
            if ((c < 0x20 && c != '\t' && c != '\r') || c == 0x7F) fail("control character in a comment");
            ++p_;
        }
    }
    void newline() {
        if (peek() == '\r') {
            if (peek(1) != '\n') fail("carriage return without a line feed");
            ++p_;
        }
        ++p_;
        ++line_;
    }
    // after a header or key/value: only whitespace and a comment may follow
    void end_of_line() {
        skip_ws();
        if (eof()) return;
        if (peek() == '#') skip_comment();
        if (eof()) return;
        if (peek() != '\n' && peek() != '\r') fail(std::string("unexpected character '") + peek() + "'");
        newline();
    }
    // whitespace, newlines and comments (inside arrays)
    void skip_blank() {
        while (!eof()) {
            skip_ws();
            if (peek() == '#') skip_comment();
            if (peek() == '\n' || peek() == '\r') {
                newline();
                continue;
            }
            break;
        }
    }

    std::string simple_key() {
        const char c = peek();
        if (c == '"') {
            if (peek(1) == '"' && peek(2) == '"') fail("multi-line strings cannot be keys");
            return basic_string();
        }
        if (c == '\'') {
            if (peek(1) == '\'' && peek(2) == '\'') fail("multi-line strings cannot be keys");
            return literal_string();
        }
        const std::size_t b = p_;
        while (!eof() && is_bare(peek())) ++p_;
        if (p_ == b) fail(std::string("expected a key, found '") + (eof() ? std::string("end of file") : std::string(1, c)) + "'");
        return s_.substr(b, p_ - b);
    }
    std::vector<std::string> dotted_key() {
        std::vector<std::string> k;
        k.push_back(simple_key());
        while (true) {
            skip_ws();
            if (peek() != '.') break;
            ++p_;
            skip_ws();
            k.push_back(simple_key());
        }
        return k;
    }

    // table headers --------------------------------------------------------
    Node* descend(Node* t, const std::string& k, bool from_header) {
        auto it = t->table.find(k);
        if (it == t->table.end()) {
            auto n = make(Node::Kind::Table);
            n->dotted = !from_header;
            Node* raw = n.get();
            t->table.emplace(k, std::move(n));
            return raw;
        }
        Node* n = it->second.get();
        if (n->kind == Node::Kind::Array && n->table_array) return n->array.back().get();
        if (n->kind != Node::Kind::Table) fail("key '" + k + "' is not a table");
        if (n->inline_table) fail("cannot extend inline table '" + k + "'");
        if (!from_header && n->explicit_header) fail("cannot extend table '" + k + "' with dotted keys");
        return n;
    }
    Node* header() {
        ++p_;
        const bool arr = peek() == '[';
        if (arr) ++p_;
        skip_ws();
        const std::vector<std::string> k = dotted_key();
        skip_ws();
        if (peek() != ']') fail("expected ']' to close a table header");
        ++p_;
        if (arr) {
            if (peek() != ']') fail("expected ']]' to close an array-of-tables header");
            ++p_;
        }
        Node* t = root_.get();
        for (std::size_t j = 0; j + 1 < k.size(); ++j) t = descend(t, k[j], true);
        const std::string& last = k.back();
        auto it = t->table.find(last);
        if (arr) {
            if (it == t->table.end()) {
                auto a = make(Node::Kind::Array);
                a->table_array = true;
                it = t->table.emplace(last, std::move(a)).first;
            }
            Node* a = it->second.get();
            if (a->kind != Node::Kind::Array || !a->table_array) fail("cannot redefine '" + last + "' as an array of tables");
            auto e = make(Node::Kind::Table);
            e->explicit_header = true;
            a->array.push_back(e);
            return e.get();
        }
        if (it == t->table.end()) {
            auto n = make(Node::Kind::Table);
            n->explicit_header = true;
            Node* raw = n.get();
            t->table.emplace(last, std::move(n));
            return raw;
        }
        Node* n = it->second.get();
        if (n->kind != Node::Kind::Table || n->inline_table) fail("cannot redefine '" + last + "' as a table");
        if (n->explicit_header || n->dotted) fail("table '" + last + "' is defined twice");
        n->explicit_header = true;
        return n;
    }

    void keyval(Node* t) {
        const std::vector<std::string> k = dotted_key();
        skip_ws();
        if (peek() != '=') fail("expected '=' after key '" + k.back() + "'");
        ++p_;
        skip_ws();
        NodePtr v = value();
        for (std::size_t j = 0; j + 1 < k.size(); ++j) t = descend(t, k[j], false);
        if (!t->table.emplace(k.back(), std::move(v)).second) fail("duplicate key '" + k.back() + "'");
    }

    // values ---------------------------------------------------------------
    NodePtr value() {
        const char c = peek();
        if (eof() || c == '\n' || c == '\r' || c == '#') fail("expected a value");
        if (c == '"' || c == '\'') {
            auto n = make(Node::Kind::String);
            if (c == '"' && peek(1) == '"' && peek(2) == '"')
                n->str = ml_basic_string();
            else if (c == '"')
                n->str = basic_string();
            else if (peek(1) == '\'' && peek(2) == '\'')
                n->str = ml_literal_string();
            else
                n->str = literal_string();
            return n;
        }
        if (c == '[') return array();
        if (c == '{') return inline_table();
        return scalar();
    }

    std::uint32_t hex_escape(int digits) {
        std::uint32_t cp = 0;
        for (int d = 0; d < digits; ++d) {
            const char h = peek();
            int v;
            if (h >= '0' && h <= '9') v = h - '0';
            else if (h >= 'a' && h <= 'f') v = h - 'a' + 10;
            else if (h >= 'A' && h <= 'F') v = h - 'A' + 10;
            else fail("invalid unicode escape");
            cp = cp * 16 + static_cast<std::uint32_t>(v);
            ++p_;
        }
        if (cp > 0x10FFFF || (cp >= 0xD800 && cp <= 0xDFFF)) fail("invalid unicode scalar in escape");
        return cp;
    }
    void escape(std::string& out) {
        ++p_;  // backslash
        const char e = peek();
        ++p_;
        switch (e) {
            case 'b': out += '\b'; break;
            case 't': out += '\t'; break;
            case 'n': out += '\n'; break;
            case 'f': out += '\f'; break;
            case 'r': out += '\r'; break;
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case 'u': append_utf8(out, hex_escape(4)); break;
            case 'U': append_utf8(out, hex_escape(8)); break;
            default: fail(std::string("invalid escape sequence '\\") + e + "'");
        }
    }
    void string_char(std::string& out) {
        const unsigned char c = static_cast<unsigned char>(peek());
        if ((c < 0x20 && c != '\t') || c == 0x7F) fail("control character in a string");
        out += static_cast<char>(c);
        ++p_;
    }
    std::string basic_string() {
        ++p_;
        std::string out;
        while (true) {
            if (eof() || peek() == '\n' || peek() == '\r') fail("unterminated string");
            if (peek() == '"') {
                ++p_;
                return out;
            }
            if (peek() == '\\')
                escape(out);
            else
                string_char(out);
        }
    }
    std::string literal_string() {
        ++p_;
        std::string out;
        while (true) {
            if (eof() || peek() == '\n' || peek() == '\r') fail("unterminated string");
            if (peek() == '\'') {
                ++p_;
                return out;
            }
            string_char(out);
        }
    }
    std::string ml_basic_string() {
        p_ += 3;
        if (peek() == '\n' || peek() == '\r') newline();
        std::string out;
        while (true) {
            if (eof()) fail("unterminated multi-line string");
            if (peek() == '"' && peek(1) == '"' && peek(2) == '"') {
                p_ += 3;
                for (int extra = 0; extra < 2 && peek() == '"'; ++extra, ++p_) out += '"';
                return out;
            }
            if (peek() == '\\') {
                // line-ending backslash: trim whitespace and newlines up to the next text
                std::size_t q = p_ + 1;
                while (q < s_.size() && (s_[q] == ' ' || s_[q] == '\t')) ++q;
                if (q < s_.size() && (s_[q] == '\n' || s_[q] == '\r')) {
                    p_ = q;
                    while (!eof() && (peek() == ' ' || peek() == '\t' || peek() == '\n' || peek() == '\r')) {
                        if (peek() == '\n' || peek() == '\r')
                            newline();
                        else
                            ++p_;
                    }
                    continue;
                }
                escape(out);
                continue;
            }
            if (peek() == '\n' || peek() == '\r') {
                newline();
                out += '\n';
                continue;
            }
            string_char(out);
        }
    }
    std::string ml_literal_string() {
        p_ += 3;
        if (peek() == '\n' || peek() == '\r') newline();
        std::string out;
        while (true) {
            if (eof()) fail("unterminated multi-line string");
            if (peek() == '\'' && peek(1) == '\'' && peek(2) == '\'') {
                p_ += 3;
                for (int extra = 0; extra < 2 && peek() == '\''; ++extra, ++p_) out += '\'';
                return out;
            }
            if (peek() == '\n' || peek() == '\r') {
                newline();
                out += '\n';
                continue;
            }
            string_char(out);
        }
    }

    NodePtr array() {
        ++p_;
        auto a = make(Node::Kind::Array);
        while (true) {
            skip_blank();
            if (peek() == ']') {
                ++p_;
                return a;
            }
            a->array.push_back(value());
            skip_blank();
            if (peek() == ',') {
                ++p_;
                continue;
            }
            if (peek() == ']') {
                ++p_;
                return a;
            }
            fail("expected ',' or ']' in an array");
        }
    }

    NodePtr inline_table() {
        ++p_;
        auto t = make(Node::Kind::Table);
        skip_ws();
        if (peek() == '}') {
            ++p_;
            t->inline_table = true;
            return t;
        }
        while (true) {
            skip_ws();
            keyval(t.get());
            skip_ws();
            if (peek() == ',') {
                ++p_;
                continue;
            }
            if (peek() == '}') {
                ++p_;
                break;
            }
            fail("expected ',' or '}' in an inline table");
        }
        t->inline_table = true;
        return t;
    }

    // digits with single underscores between them
    static bool digits_ok(const std::string& d, bool (*ok)(char)) {
        if (d.empty() || d.front() == '_' || d.back() == '_') return false;
        for (std::size_t k = 0; k < d.size(); ++k) {
            if (d[k] == '_') {
                if (d[k + 1] == '_') return false;
            } else if (!ok(d[k])) {
                return false;
            }
        }
        return true;
    }
    static std::string strip(const std::string& d) {
        std::string o;
        for (char c : d)
            if (c != '_') o += c;
        return o;
    }

    // two digits at s[k], as a number
    static int two(const std::string& s, std::size_t k) {
        if (k + 1 >= s.size() || !is_digit(s[k]) || !is_digit(s[k + 1])) return -1;
        return (s[k] - '0') * 10 + (s[k + 1] - '0');
    }
    // HH:MM:SS[.frac] at s[k]; returns the index after it, or npos
    static std::size_t time_part(const std::string& s, std::size_t k) {
        const int h = two(s, k), m = two(s, k + 3), sec = two(s, k + 6);
        if (h < 0 || h > 23 || k + 2 >= s.size() || s[k + 2] != ':' || m < 0 || m > 59 || k + 5 >= s.size() ||
            s[k + 5] != ':' || sec < 0 || sec > 59)
            return std::string::npos;
        std::size_t e = k + 8;
        if (e < s.size() && s[e] == '.') {
            const std::size_t f = ++e;
            while (e < s.size() && is_digit(s[e])) ++e;
            if (e == f) return std::string::npos;
        }
        return e;
    }
    // local date, local / offset date-time, local time (TOML 1.0), kept as written
    NodePtr date_time(const std::string& t, const std::string& shown) {
        std::size_t e = 0;
        if (t.size() >= 5 && t[4] == '-') {
            const int y = (t[0] - '0') * 1000 + (t[1] - '0') * 100 + (t[2] - '0') * 10 + (t[3] - '0');
            const int mo = two(t, 5), d = two(t, 8);
            static const int kDays[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
            const bool leap = (y % 4 == 0 && y % 100 != 0) || y % 400 == 0;
            if (t.size() < 10 || t[7] != '-' || mo < 1 || mo > 12 || d < 1 ||
                d > kDays[mo - 1] + ((mo == 2 && leap) ? 1 : 0))
                fail("invalid date '" + shown + "'");
            e = 10;
            if (e < t.size()) {
                if (t[e] != 'T' && t[e] != 't') fail("invalid date-time '" + shown + "'");
                e = time_part(t, e + 1);
                if (e == std::string::npos) fail("invalid date-time '" + shown + "'");
                if (e < t.size()) {  // offset
                    if (t[e] == 'Z' || t[e] == 'z') {
                        ++e;
                    } else if (t[e] == '+' || t[e] == '-') {
                        const int oh = two(t, e + 1), om = two(t, e + 4);
                        if (oh < 0 || oh > 23 || e + 3 >= t.size() || t[e + 3] != ':' || om < 0 || om > 59)
                            fail("invalid date-time offset '" + shown + "'");
                        e += 6;
                    }
                }
            }
        } else {
            e = time_part(t, 0);
            if (e == std::string::npos) fail("invalid time '" + shown + "'");
        }
        if (e != t.size()) fail("invalid date or time '" + shown + "'");
        auto n = make(Node::Kind::DateTime);
        n->str = t;
        return n;
    }

    NodePtr scalar() {
        const std::size_t b = p_;
        while (!eof()) {
            const char c = peek();
            if (c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == ',' || c == ']' || c == '}' || c == '#') break;
            ++p_;
        }
        const std::string tok = s_.substr(b, p_ - b);
        if (tok == "true" || tok == "false") {
            auto n = make(Node::Kind::Boolean);
            n->b = tok == "true";
            return n;
        }
        if (tok == "inf" || tok == "+inf" || tok == "-inf" || tok == "nan" || tok == "+nan" || tok == "-nan") {
            auto n = make(Node::Kind::Float);
            const double v = tok.back() == 'f' ? std::numeric_limits<double>::infinity()
                                               : std::numeric_limits<double>::quiet_NaN();
            n->f = tok[0] == '-' ? -v : v;
            return n;
        }
        if (tok.size() >= 5 && is_digit(tok[0]) && is_digit(tok[1]) && is_digit(tok[2]) && is_digit(tok[3]) &&
            tok[4] == '-') {
            std::string dt = tok;
            // "date time" with a space separator: the time is the next token
            if (tok.size() == 10 && peek() == ' ' && is_digit(peek(1)) && is_digit(peek(2)) && peek(3) == ':') {
                const std::size_t b2 = ++p_;
                while (!eof()) {
                    const char c = peek();
                    if (c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == ',' || c == ']' || c == '}' || c == '#')
                        break;
                    ++p_;
                }
                dt += "T" + s_.substr(b2, p_ - b2);
            }
            return date_time(dt, tok);
        }
        if (tok.size() >= 3 && is_digit(tok[0]) && is_digit(tok[1]) && tok[2] == ':') return date_time(tok, tok);
        if (tok.size() > 2 && tok[0] == '0' && (tok[1] == 'x' || tok[1] == 'o' || tok[1] == 'b')) {
            const int base = tok[1] == 'x' ? 16 : tok[1] == 'o' ? 8 : 2;
            const std::string d = tok.substr(2);
            bool (*ok)(char) = base == 16 ? +[](char c) {
                return is_digit(c) || (c >= 'a' && c <= 'f') || (c >= 'A' && c <= 'F');
            } : base == 8 ? +[](char c) { return c >= '0' && c <= '7'; } : +[](char c) { return c == '0' || c == '1'; };
            if (!digits_ok(d, ok)) fail("invalid integer '" + tok + "'");
            errno = 0;
            char* end = nullptr;
            const std::string clean = strip(d);
            const unsigned long long v = std::strtoull(clean.c_str(), &end, base);
            if (errno == ERANGE || v > static_cast<unsigned long long>(std::numeric_limits<std::int64_t>::max()))
                fail("integer '" + tok + "' out of range");
            auto n = make(Node::Kind::Integer);
            n->i = static_cast<std::int64_t>(v);
            return n;
        }
        std::string body = tok;
        if (!body.empty() && (body[0] == '+' || body[0] == '-')) body = body.substr(1);
        const std::size_t dot = body.find('.');
        const std::size_t ex = body.find_first_of("eE");
        const std::string ipart = body.substr(0, std::min(dot, ex));
        if (!digits_ok(ipart, is_digit)) fail("invalid value '" + tok + "'");
        if (ipart.size() > 1 && ipart[0] == '0') fail("leading zeros are not allowed in '" + tok + "'");
        if (dot == std::string::npos && ex == std::string::npos) {
            errno = 0;
            const std::string clean = strip(tok);
            char* end = nullptr;
            const long long v = std::strtoll(clean.c_str(), &end, 10);
            if (errno == ERANGE) fail("integer '" + tok + "' out of range");
            auto n = make(Node::Kind::Integer);
            n->i = v;
            return n;
        }
        if (dot != std::string::npos) {
            if (ex != std::string::npos && ex < dot) fail("invalid float '" + tok + "'");
            const std::string frac = body.substr(dot + 1, (ex == std::string::npos ? body.size() : ex) - dot - 1);
            if (!digits_ok(frac, is_digit)) fail("invalid float '" + tok + "'");
        }
        if (ex != std::string::npos) {
            std::string e = body.substr(ex + 1);
            if (!e.empty() && (e[0] == '+' || e[0] == '-')) e = e.substr(1);
            if (!digits_ok(e, is_digit)) fail("invalid float '" + tok + "'");
        }
        const std::string clean = strip(tok);
        char* end = nullptr;
        const double v = std::strtod(clean.c_str(), &end);
        if (end != clean.c_str() + clean.size()) fail("invalid float '" + tok + "'");
        auto n = make(Node::Kind::Float);
        n->f = v;
        return n;
    }
};

}  // namespace

NodePtr parse(const std::string& text) { return Parser(text).run(); }

}  // namespace cfdg::toml
