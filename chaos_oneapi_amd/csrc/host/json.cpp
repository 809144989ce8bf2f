#include "json.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace chaos {

namespace {

struct Parser {
    const std::string& s;
    size_t p = 0;

    explicit Parser(const std::string& text) : s(text) {}

    [[noreturn]] void error(const char* what) const {
        throw std::invalid_argument(std::string("json parse error at offset ") +
                                    std::to_string(p) + ": " + what);
    }
    void ws() {
        while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r')) ++p;
    }
    bool eat(char c) {
        ws();
        if (p < s.size() && s[p] == c) {
            ++p;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) error((std::string("expected '") + c + "'").c_str());
    }
    std::string str() {
        expect('"');
        std::string out;
        while (true) {
            if (p >= s.size()) error("unterminated string");
            char c = s[p++];
            if (c == '"') break;
            if (c == '\\') {
                if (p >= s.size()) error("bad escape");
                char e = s[p++];
                switch (e) {
                    case '"': out += '"'; break;
                    case '\\': out += '\\'; break;
                    case '/': out += '/'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'n': out += '\n'; break;
                    case 'r': out += '\r'; break;
                    case 't': out += '\t'; break;
                    case 'u': {
                        if (p + 4 > s.size()) error("bad \\u escape");
                        unsigned cp = (unsigned)std::strtoul(s.substr(p, 4).c_str(), nullptr, 16);
                        p += 4;
                        if (cp < 0x80) {
                            out += (char)cp;
                        } else if (cp < 0x800) {
                            out += (char)(0xC0 | (cp >> 6));
                            out += (char)(0x80 | (cp & 0x3F));
                        } else {
                            out += (char)(0xE0 | (cp >> 12));
                            out += (char)(0x80 | ((cp >> 6) & 0x3F));
                            out += (char)(0x80 | (cp & 0x3F));
                        }
                        break;
                    }
                    default: error("bad escape");
                }
            } else {
                out += c;
            }
        }
        return out;
    }
    Json value() {
        ws();
        if (p >= s.size()) error("unexpected end");
        char c = s[p];
        if (c == '{') {
            ++p;
            Json o = Json::object();
            if (eat('}')) return o;
            do {
                ws();
                std::string k = str();
                expect(':');
                o[k] = value();  // a repeated key keeps the last value
            } while (eat(','));
            expect('}');
            return o;
        }
        if (c == '[') {
            ++p;
            Json a = Json::array();
            if (eat(']')) return a;
            do {
                a.push_back(value());
            } while (eat(','));
            expect(']');
            return a;
        }
        if (c == '"') return Json(str());
        if (s.compare(p, 4, "true") == 0) {
            p += 4;
            return Json(true);
        }
        if (s.compare(p, 5, "false") == 0) {
            p += 5;
            return Json(false);
        }
        if (s.compare(p, 4, "null") == 0) {
            p += 4;
            return Json();
        }
        size_t q = p;
        bool is_float = false;
        if (q < s.size() && (s[q] == '-' || s[q] == '+')) ++q;
        while (q < s.size() && (std::isdigit((unsigned char)s[q]) || s[q] == '.' || s[q] == 'e' ||
                                s[q] == 'E' || s[q] == '-' || s[q] == '+')) {
            if (s[q] == '.' || s[q] == 'e' || s[q] == 'E') is_float = true;
            ++q;
        }
        if (q == p) error("unexpected character");
        std::string num = s.substr(p, q - p);
        p = q;
        char* end = nullptr;
        if (!is_float) {
            long long v = std::strtoll(num.c_str(), &end, 10);
            if (*end == 0) return Json((int64_t)v);
        }
        double d = std::strtod(num.c_str(), &end);
        if (*end != 0) error("bad number");
        return Json(d);
    }
};

void escape_into(std::string& out, const std::string& s) {
    out += '"';
    for (char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            case '\b': out += "\\b"; break;
            case '\f': out += "\\f"; break;
            default:
                if ((unsigned char)c < 0x20) {
                    char buf[8];
                    std::snprintf(buf, sizeof(buf), "\\u%04x", (unsigned)(unsigned char)c);
                    out += buf;
                } else {
                    out += c;
                }
        }
    }
    out += '"';
}

}  // namespace

std::string format_double(double v) {
    if (std::isnan(v) || std::isinf(v)) return "null";  // JSON has no NaN/Inf
    if (v == 0.0) return std::signbit(v) ? "-0.0" : "0.0";
    // shortest round-trip significand digits and decimal exponent
    char buf[40];
    int prec = 1;
    for (; prec <= 17; ++prec) {
        std::snprintf(buf, sizeof(buf), "%.*e", prec - 1, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    std::string sci(buf);
    const bool neg = sci[0] == '-';
    const size_t epos = sci.find('e');
    std::string digits;
    for (size_t i = neg ? 1 : 0; i < epos; ++i)
        if (sci[i] != '.') digits += sci[i];
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
    const int e10 = std::atoi(sci.c_str() + epos + 1);
    const int k = (int)digits.size();
    const int n = e10 + 1;  // value = 0.d1..dk * 10^n
    // layout of nlohmann/json's dtoa format_buffer (min_exp -4, max_exp 15)
    std::string out = neg ? "-" : "";
    if (k <= n && n <= 15) {
        out += digits + std::string((size_t)(n - k), '0') + ".0";
    } else if (0 < n && n <= 15) {
        out += digits.substr(0, (size_t)n) + "." + digits.substr((size_t)n);
    } else if (-4 < n && n <= 0) {
        out += "0." + std::string((size_t)(-n), '0') + digits;
    } else {
        out += digits.substr(0, 1);
        if (k > 1) out += "." + digits.substr(1);
        char eb[16];
        std::snprintf(eb, sizeof(eb), "e%c%02d", n - 1 < 0 ? '-' : '+', std::abs(n - 1));
        out += eb;
    }
    return out;
}

Json::Json(const std::vector<double>& v) : type_(Type::Array) {
    for (double x : v) a_.emplace_back(x);
}
Json::Json(const std::vector<int>& v) : type_(Type::Array) {
    for (int x : v) a_.emplace_back(x);
}

Json Json::array() {
    Json j;
    j.type_ = Type::Array;
    return j;
}
Json Json::object() {
    Json j;
    j.type_ = Type::Object;
    return j;
}

Json Json::parse(const std::string& text) {
    Parser ps(text);
    Json v = ps.value();
    ps.ws();
    if (ps.p != text.size()) ps.error("trailing characters");
    return v;
}

const Json& Json::at(const std::string& key) const {
    if (type_ != Type::Object) throw std::out_of_range("json: not an object (key '" + key + "')");
    auto it = o_.find(key);
    if (it == o_.end()) throw std::out_of_range("json: key '" + key + "' not found");
    return it->second;
}

bool Json::contains(const std::string& key) const {
    return type_ == Type::Object && o_.count(key) > 0;
}

Json& Json::operator[](const std::string& key) {
    if (type_ == Type::Null) type_ = Type::Object;
    if (type_ != Type::Object) throw std::invalid_argument("json: operator[] on a non-object");
    return o_[key];
}

const std::map<std::string, Json>& Json::items() const {
    if (type_ != Type::Object) throw std::invalid_argument("json: items() on a non-object");
    return o_;
}

size_t Json::size() const {
    if (type_ == Type::Array) return a_.size();
    if (type_ == Type::Object) return o_.size();
    return 0;
}

const Json& Json::at(size_t i) const {
    if (type_ != Type::Array) throw std::out_of_range("json: not an array");
    return a_.at(i);
}

void Json::push_back(const Json& v) {
    if (type_ == Type::Null) type_ = Type::Array;
    if (type_ != Type::Array) throw std::invalid_argument("json: push_back on a non-array");
    a_.push_back(v);
}

double Json::as_double() const {
    if (type_ == Type::Double) return d_;
    if (type_ == Type::Int) return (double)i_;
    throw std::invalid_argument("json: not a number");
}

int64_t Json::as_int() const {
    if (type_ == Type::Int) return i_;
    if (type_ == Type::Double) return (int64_t)d_;
    throw std::invalid_argument("json: not a number");
}

bool Json::as_bool() const {
    if (type_ == Type::Bool) return b_;
    throw std::invalid_argument("json: not a boolean");
}

const std::string& Json::as_string() const {
    if (type_ == Type::String) return s_;
    throw std::invalid_argument("json: not a string");
}

std::vector<double> Json::as_double_vector() const {
    if (type_ != Type::Array) throw std::invalid_argument("json: not an array");
    std::vector<double> v;
    v.reserve(a_.size());
    for (const auto& x : a_) v.push_back(x.as_double());
    return v;
}

std::string Json::dump(int indent) const {
    std::string out;
    dump_to(out, indent, 0);
    return out;
}

void Json::dump_to(std::string& out, int indent, int level) const {
    auto newline = [&](int lvl) {
        if (indent >= 0) {
            out += '\n';
            out.append((size_t)(indent * lvl), ' ');
        }
    };
    switch (type_) {
        case Type::Null: out += "null"; break;
        case Type::Bool: out += b_ ? "true" : "false"; break;
        case Type::Int: out += std::to_string(i_); break;
        case Type::Double: out += format_double(d_); break;
        case Type::String: escape_into(out, s_); break;
        case Type::Array: {
            if (a_.empty()) {
                out += "[]";
                break;
            }
            out += '[';
            for (size_t i = 0; i < a_.size(); ++i) {
                newline(level + 1);
                a_[i].dump_to(out, indent, level + 1);
                if (i + 1 < a_.size()) out += ',';
            }
            newline(level);
            out += ']';
            break;
        }
        case Type::Object: {
            if (o_.empty()) {
                out += "{}";
                break;
            }
            out += '{';
            size_t i = 0;
            for (const auto& kv : o_) {
                newline(level + 1);
                escape_into(out, kv.first);
                out += indent >= 0 ? ": " : ":";
                kv.second.dump_to(out, indent, level + 1);
                if (++i < o_.size()) out += ',';
            }
            newline(level);
            out += '}';
            break;
        }
    }
}

}  // namespace chaos
