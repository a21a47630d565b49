// A parsed configuration document.
//
// Accepts what the reference's parameter files contain (its loader:
// src/json.cpp, include/astrosimgpu/json.hpp): JSON objects, arrays,
// numbers, strings, true/false/null, and "//" comments running to the end of
// a line. A later duplicate key replaces an earlier one. Every lookup of an
// object member is remembered, so that the loader can reject keys no one
// asked for -- the reference's guard against typos and options a binary does
// not know (parameters.cpp:191-202).
//
// Layout: every value of the document is one entry of a flat arena; an
// object or array entry lists the arena indices of its members. Parsing is
// iterative (an explicit stack of open containers), so nesting depth costs
// heap, not call stack. Node is a cheap handle (document + index) that
// member lookups return; a missing member is the null handle.
#pragma once

#include <cstdint>
#include <string>
#include <type_traits>
#include <vector>

namespace asg_host {

class ConfigDoc {
public:
    enum class Kind : uint8_t { Null, Bool, Number, String, Array, Object };

    class Node {
    public:
        Node() = default;
        Kind kind() const { return doc_ ? doc_->values_[id_].kind : Kind::Null; }
        bool is_object() const { return kind() == Kind::Object; }
        double number() const { return doc_->values_[id_].num; }

        // The member `key` of an object (the null handle if absent or if
        // this is not an object); the lookup is remembered.
        Node at(const std::string& key) const;

        // Copies member `key` into `out` when it holds a value of the
        // matching kind (a number for arithmetic types, true/false for bool,
        // a string for std::string); otherwise leaves `out` as it was.
        template <typename T>
        void take(const std::string& key, T& out) const {
            const Node v = at(key);
            if constexpr (std::is_same_v<T, bool>) {
                if (v.kind() == Kind::Bool) out = v.doc_->values_[v.id_].flag;
            } else if constexpr (std::is_arithmetic_v<T>) {
                if (v.kind() == Kind::Number) out = static_cast<T>(v.number());
            } else if constexpr (std::is_same_v<T, std::string>) {
                if (v.kind() == Kind::String) out = v.doc_->values_[v.id_].text;
            } else {
                static_assert(std::is_same_v<T, void>, "unsupported member type");
            }
        }

    private:
        friend class ConfigDoc;
        Node(const ConfigDoc* doc, uint32_t id) : doc_(doc), id_(id) {}
        const ConfigDoc* doc_ = nullptr;
        uint32_t id_ = 0;
    };

    // Throw std::runtime_error on malformed text ("JSON parse error at byte
    // N: ...") or an unreadable file ("cannot open configuration file: P").
    static ConfigDoc parse(const std::string& text);
    static ConfigDoc load(const std::string& path);

    Node root() const { return Node(this, root_); }

    // Sets the member at a dotted path ("a.b.c") to the JSON literal
    // `literal`, creating objects along the path (a non-object on the path
    // is replaced by one).
    void assign(const std::string& dotted_path, const std::string& literal);

    // Dotted paths of the members no lookup asked for; the members of a
    // member that was asked for are listed themselves. Keys are visited in
    // byte order at each level.
    std::vector<std::string> unread_paths() const;

private:
    struct Value {
        Kind kind = Kind::Null;
        bool flag = false;
        double num = 0.0;
        std::string text;
        std::vector<uint32_t> slots;  // object members / array items (indices into slots_)
    };
    struct Slot {
        std::string key;             // empty for array items
        uint32_t value = 0;
        mutable bool asked = false;
    };

    uint32_t parse_into(const std::string& text);  // appends a value tree, returns its root
    uint32_t add_value(Kind kind);
    uint32_t member_slot(uint32_t object, const std::string& key);  // finds or appends
    void unread_of(uint32_t value, const std::string& prefix, std::vector<std::string>& out) const;

    std::vector<Value> values_;
    std::vector<Slot> slots_;
    uint32_t root_ = 0;

    friend class ConfigParser;
};

}  // namespace asg_host
