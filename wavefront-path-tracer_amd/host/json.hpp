// Minimal JSON document model for scene manifests (objects, arrays, strings,
// numbers, booleans, null).  Parse errors throw std::runtime_error.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace wpt {

class Json {
public:
    enum class Type { Null, Bool, Number, String, Array, Object };

    Json() = default;
    static Json parse(const std::string& text);

    Type type() const { return type_; }
    bool isObject() const { return type_ == Type::Object; }
    bool isArray() const { return type_ == Type::Array; }
    bool isString() const { return type_ == Type::String; }
    bool isNumber() const { return type_ == Type::Number; }
    bool isBool() const { return type_ == Type::Bool; }

    bool contains(const std::string& key) const { return isObject() && object_.count(key) != 0; }
    const Json& at(const std::string& key) const;
    const std::vector<Json>& items() const { return array_; }
    const std::map<std::string, Json>& members() const { return object_; }

    const std::string& str() const;
    double number() const;
    bool boolean() const;

    // Typed lookups with defaults; a present key of the wrong type throws.
    std::string value(const std::string& key, const std::string& fallback) const;
    double value(const std::string& key, double fallback) const;
    bool value(const std::string& key, bool fallback) const;
    int value(const std::string& key, int fallback) const;

private:
    friend class JsonParser;
    Type type_ = Type::Null;
    bool bool_ = false;
    double number_ = 0.0;
    std::string string_;
    std::vector<Json> array_;
    std::map<std::string, Json> object_;
};

}  // namespace wpt
