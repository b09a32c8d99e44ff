// Minimal recursive-descent JSON reader. See json.hpp.
#include "json.hpp"

#include <cctype>
#include <cstdlib>

namespace dg {

namespace {

struct Reader {
    const std::string& s;
    size_t i = 0;

    [[noreturn]] void fail(const std::string& what) const {
        throw std::runtime_error("JSON parse error at offset " + std::to_string(i) + ": " + what);
    }
    void ws() {
        while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i]))) ++i;
    }
    bool eat(char c) {
        ws();
        if (i < s.size() && s[i] == c) {
            ++i;
            return true;
        }
        return false;
    }
    std::string str() {
        if (!eat('"')) fail("expected string");
        std::string out;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\' && i < s.size()) {
                char e = s[i++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (i + 4 > s.size()) fail("bad escape");
                        unsigned cp = std::strtoul(s.substr(i, 4).c_str(), nullptr, 16);
                        i += 4;
                        if (cp < 0x80) {
                            out += static_cast<char>(cp);
                        } else if (cp < 0x800) {
                            out += static_cast<char>(0xC0 | (cp >> 6));
                            out += static_cast<char>(0x80 | (cp & 0x3F));
                        } else {
                            out += static_cast<char>(0xE0 | (cp >> 12));
                            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
                            out += static_cast<char>(0x80 | (cp & 0x3F));
                        }
                        break;
                    }
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i >= s.size()) fail("unterminated string");
        ++i;
        return out;
    }
    JsonValue value() {
        ws();
        if (i >= s.size()) fail("unexpected end");
        JsonValue v;
        const char c = s[i];
        if (c == '{') {
            ++i;
            v.kind = JsonValue::Object;
            if (eat('}')) return v;
            do {
                std::string k = str();
                if (!eat(':')) fail("expected ':'");
                v.obj.emplace_back(k, value());
            } while (eat(','));
            if (!eat('}')) fail("expected '}'");
        } else if (c == '[') {
            ++i;
            v.kind = JsonValue::Array;
            if (eat(']')) return v;
            do {
                v.arr.push_back(value());
            } while (eat(','));
            if (!eat(']')) fail("expected ']'");
        } else if (c == '"') {
            v.kind = JsonValue::String;
            v.str = str();
        } else if (s.compare(i, 4, "true") == 0) {
            i += 4;
            v.kind = JsonValue::Bool;
            v.b = true;
        } else if (s.compare(i, 5, "false") == 0) {
            i += 5;
            v.kind = JsonValue::Bool;
        } else if (s.compare(i, 4, "null") == 0) {
            i += 4;
        } else {
            char* end = nullptr;
            v.num = std::strtod(s.c_str() + i, &end);
            if (end == s.c_str() + i) fail("bad value");
            i = end - s.c_str();
            v.kind = JsonValue::Number;
        }
        return v;
    }
};

}  // namespace

JsonValue parseJson(const std::string& text) {
    Reader r{text};
    JsonValue v = r.value();
    r.ws();
    if (r.i != text.size()) r.fail("trailing characters");
    return v;
}

}  // namespace dg
