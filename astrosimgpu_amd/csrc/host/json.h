// Minimal read-only JSON for the configuration files: objects, arrays,
// numbers, strings, booleans, null, and // line comments (which the
// reference's config files use, json.cpp:45-50). Every object remembers which
// keys were looked up so a loader can reject keys nobody read, as the
// reference does (json.cpp:258-269, parameters.cpp:191-202).
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace asg_host {

class JsonValue {
public:
    enum class Kind { Null, Bool, Number, String, Array, Object };

    static JsonValue parse(const std::string& text);  // throws std::runtime_error
    static JsonValue parse_file(const std::string& path);

    Kind kind() const { return kind_; }
    bool is_object() const { return kind_ == Kind::Object; }
    bool is_number() const { return kind_ == Kind::Number; }
    bool is_bool() const { return kind_ == Kind::Bool; }
    bool is_string() const { return kind_ == Kind::String; }
    double number() const { return num_; }
    bool boolean() const { return flag_; }
    const std::string& str() const { return str_; }

    // Object member lookup; marks the key as used. Absent -> a null value.
    const JsonValue& get(const std::string& key) const;
    // Set or replace a member (used for command-line style overrides).
    JsonValue& member(const std::string& key);
    void set_number(double v) { kind_ = Kind::Number; num_ = v; }
    void set_parsed(const JsonValue& v) { *this = v; }

    // Keys present in the file that no lookup asked for, as dotted paths.
    void unused_keys(const std::string& prefix, std::vector<std::string>& out) const;

    // Typed readers: assign only when the key holds a value of that kind.
    void read(const std::string& key, double& out) const;
    void read(const std::string& key, int& out) const;
    void read(const std::string& key, unsigned& out) const;
    void read(const std::string& key, unsigned long long& out) const;
    void read(const std::string& key, unsigned long& out) const;
    void read(const std::string& key, bool& out) const;
    void read(const std::string& key, std::string& out) const;

private:
    friend class JsonReader;
    Kind kind_ = Kind::Null;
    double num_ = 0.0;
    bool flag_ = false;
    std::string str_;
    std::vector<JsonValue> items_;
    std::vector<std::pair<std::string, JsonValue>> members_;  // file order
    mutable std::vector<bool> used_;
};

}  // namespace asg_host
