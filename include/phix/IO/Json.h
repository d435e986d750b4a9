#pragma once
// A JSON value and parser for the configuration layer (the reference reads its
// settings through nlohmann::json, include/IO/ConfigFile.h:3, which this image
// does not have). It offers the part of that value interface the reference's
// drivers use: operator[] / at / contains on objects, get<T>(), value(key,
// default), and implicit conversion to arithmetic types and std::string, so
// `const int nx = cfg["mesh"]["nx"];` reads as it does there.
//
// parse() accepts standard JSON plus `//` line comments outside strings
// (JSONC, as the reference's ConfigFile strips them, src/IO/ConfigFile.cpp).
// Errors throw std::runtime_error (parse, with line:column) or
// std::out_of_range (missing key or index) / std::runtime_error (type).
#include <cstddef>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace PhiX {
namespace IO {

class Json {
public:
    enum class Type { Null, Boolean, Number, String, Array, Object };

    Json() = default;
    static Json parse(const std::string& text);

    Type type() const { return type_; }
    bool is_null() const { return type_ == Type::Null; }
    bool is_boolean() const { return type_ == Type::Boolean; }
    bool is_number() const { return type_ == Type::Number; }
    bool is_number_integer() const { return type_ == Type::Number && integral_; }
    bool is_string() const { return type_ == Type::String; }
    bool is_array() const { return type_ == Type::Array; }
    bool is_object() const { return type_ == Type::Object; }

    // objects: members in file order; a repeated key keeps its last value
    const Json& operator[](const std::string& key) const { return at(key); }
    const Json& operator[](const char* key) const { return at(std::string(key)); }
    const Json& at(const std::string& key) const;
    bool contains(const std::string& key) const;
    const std::vector<std::pair<std::string, Json>>& items() const;

    // arrays
    const Json& operator[](std::size_t i) const { return at(i); }
    const Json& operator[](int i) const { return at((std::size_t)i); }
    const Json& at(std::size_t i) const;
    // elements of an array, members of an object, 0 for null, 1 otherwise
    std::size_t size() const;

    template <class T>
    T get() const;

    template <class T>
    T value(const std::string& key, const T& def) const {
        return contains(key) ? at(key).get<T>() : def;
    }
    std::string value(const std::string& key, const char* def) const;

    template <class T, typename = std::enable_if_t<std::is_arithmetic<T>::value || std::is_same<T, std::string>::value>>
    operator T() const {
        return get<T>();
    }

    // compact serialisation (numbers with 17 significant digits)
    std::string dump() const;

private:
    friend class JsonParser;
    Type type_ = Type::Null;
    bool bool_ = false;
    bool integral_ = false;   // written without fraction or exponent
    double num_ = 0.0;
    long long int_ = 0;       // exact value when integral_
    std::string str_;
    std::vector<Json> arr_;
    std::vector<std::pair<std::string, Json>> obj_;

    [[noreturn]] void typeError(const char* want) const;
    double asDouble() const;
    long long asInteger() const;
    const std::string& asString() const;
    bool asBool() const;
};

template <> inline bool Json::get<bool>() const { return asBool(); }
template <> inline int Json::get<int>() const { return (int)asInteger(); }
template <> inline long Json::get<long>() const { return (long)asInteger(); }
template <> inline long long Json::get<long long>() const { return asInteger(); }
template <> inline unsigned Json::get<unsigned>() const { return (unsigned)asInteger(); }
template <> inline std::size_t Json::get<std::size_t>() const { return (std::size_t)asInteger(); }
template <> inline float Json::get<float>() const { return (float)asDouble(); }
template <> inline double Json::get<double>() const { return asDouble(); }
template <> inline std::string Json::get<std::string>() const { return asString(); }

inline std::string Json::value(const std::string& key, const char* def) const {
    return contains(key) ? at(key).get<std::string>() : std::string(def);
}

}  // namespace IO
}  // namespace PhiX
