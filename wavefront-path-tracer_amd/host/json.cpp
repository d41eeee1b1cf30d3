#include "json.hpp"

#include <cctype>
#include <cstdlib>
#include <stdexcept>

namespace wpt {

class JsonParser {
public:
    explicit JsonParser(const std::string& t) : text(t) {}

    Json document() {
        Json v = value();
        skip();
        if (pos != text.size()) error("trailing characters");
        return v;
    }

private:
    const std::string& text;
    size_t pos = 0;

    [[noreturn]] void error(const char* what) {
        throw std::runtime_error(std::string("JSON ") + what + " at offset " + std::to_string(pos));
    }
    void skip() {
        while (pos < text.size() && std::isspace(static_cast<unsigned char>(text[pos]))) ++pos;
    }
    bool consume(char c) {
        skip();
        if (pos < text.size() && text[pos] == c) {
            ++pos;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!consume(c)) error("unexpected character");
    }
    bool literal(const char* word) {
        size_t n = std::char_traits<char>::length(word);
        if (text.compare(pos, n, word) == 0) {
            pos += n;
            return true;
        }
        return false;
    }
    std::string string() {
        expect('"');
        std::string out;
        while (true) {
            if (pos >= text.size()) error("unterminated string");
            char c = text[pos++];
            if (c == '"') break;
            if (c != '\\') {
                out.push_back(c);
                continue;
            }
            if (pos >= text.size()) error("bad escape");
            char e = text[pos++];
            switch (e) {
            case '"': out.push_back('"'); break;
            case '\\': out.push_back('\\'); break;
            case '/': out.push_back('/'); break;
            case 'b': out.push_back('\b'); break;
            case 'f': out.push_back('\f'); break;
            case 'n': out.push_back('\n'); break;
            case 'r': out.push_back('\r'); break;
            case 't': out.push_back('\t'); break;
            case 'u': {
                if (pos + 4 > text.size()) error("bad unicode escape");
                unsigned code = static_cast<unsigned>(std::strtoul(text.substr(pos, 4).c_str(), nullptr, 16));
                pos += 4;
                if (code < 0x80) {
                    out.push_back(static_cast<char>(code));
                } else if (code < 0x800) {
                    out.push_back(static_cast<char>(0xc0 | (code >> 6)));
                    out.push_back(static_cast<char>(0x80 | (code & 0x3f)));
                } else {
                    out.push_back(static_cast<char>(0xe0 | (code >> 12)));
                    out.push_back(static_cast<char>(0x80 | ((code >> 6) & 0x3f)));
                    out.push_back(static_cast<char>(0x80 | (code & 0x3f)));
                }
                break;
            }
            default: error("bad escape");
            }
        }
        return out;
    }
    Json value() {
        skip();
        if (pos >= text.size()) error("unexpected end");
        Json v;
        char c = text[pos];
        if (c == '{') {
            ++pos;
            v.type_ = Json::Type::Object;
            if (consume('}')) return v;
            do {
                skip();
                std::string key = string();
                expect(':');
                v.object_[key] = value();
            } while (consume(','));
            expect('}');
        } else if (c == '[') {
            ++pos;
            v.type_ = Json::Type::Array;
            if (consume(']')) return v;
            do {
                v.array_.push_back(value());
            } while (consume(','));
            expect(']');
        } else if (c == '"') {
            v.type_ = Json::Type::String;
            v.string_ = string();
        } else if (literal("true")) {
            v.type_ = Json::Type::Bool;
            v.bool_ = true;
        } else if (literal("false")) {
            v.type_ = Json::Type::Bool;
            v.bool_ = false;
        } else if (literal("null")) {
            v.type_ = Json::Type::Null;
        } else if (c == '-' || std::isdigit(static_cast<unsigned char>(c))) {
            const char* begin = text.c_str() + pos;
            char* end = nullptr;
            double d = std::strtod(begin, &end);
            if (end == begin) error("bad number");
            pos += static_cast<size_t>(end - begin);
            v.type_ = Json::Type::Number;
            v.number_ = d;
        } else {
            error("unexpected character");
        }
        return v;
    }
};

Json Json::parse(const std::string& text) { return JsonParser(text).document(); }

const Json& Json::at(const std::string& key) const {
    auto it = object_.find(key);
    if (!isObject() || it == object_.end()) throw std::runtime_error("missing key: " + key);
    return it->second;
}

const std::string& Json::str() const {
    if (!isString()) throw std::runtime_error("expected a string");
    return string_;
}

double Json::number() const {
    if (!isNumber()) throw std::runtime_error("expected a number");
    return number_;
}

bool Json::boolean() const {
    if (!isBool()) throw std::runtime_error("expected a boolean");
    return bool_;
}

std::string Json::value(const std::string& key, const std::string& fallback) const {
    return contains(key) ? at(key).str() : fallback;
}
double Json::value(const std::string& key, double fallback) const { return contains(key) ? at(key).number() : fallback; }
bool Json::value(const std::string& key, bool fallback) const { return contains(key) ? at(key).boolean() : fallback; }
int Json::value(const std::string& key, int fallback) const {
    return contains(key) ? static_cast<int>(at(key).number()) : fallback;
}

}  // namespace wpt
