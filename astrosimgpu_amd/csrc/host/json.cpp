#include "json.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace asg_host {

namespace {
const JsonValue& null_value() {
    static const JsonValue v;
    return v;
}
}  // namespace

class JsonReader {
public:
    explicit JsonReader(const std::string& s) : s_(s) {}

    JsonValue document() {
        blank();
        JsonValue v = value();
        blank();
        if (p_ != s_.size()) error("unexpected text after the top-level value");
        return v;
    }

private:
    [[noreturn]] void error(const std::string& msg) const {
        throw std::runtime_error("JSON parse error at byte " + std::to_string(p_) + ": " + msg);
    }

    // Whitespace and // comments.
    void blank() {
        while (p_ < s_.size()) {
            const char c = s_[p_];
            if (c == ' ' || c == '\t' || c == '\r' || c == '\n') {
                ++p_;
            } else if (c == '/' && p_ + 1 < s_.size() && s_[p_ + 1] == '/') {
                while (p_ < s_.size() && s_[p_] != '\n') ++p_;
            } else {
                break;
            }
        }
    }

    char at() const {
        if (p_ >= s_.size()) throw std::runtime_error("JSON parse error: unexpected end of input");
        return s_[p_];
    }

    void want(char c) {
        if (p_ >= s_.size() || s_[p_] != c) error(std::string("expected '") + c + "'");
        ++p_;
    }

    JsonValue value() {
        const char c = at();
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') {
            JsonValue v;
            v.kind_ = JsonValue::Kind::String;
            v.str_ = string();
            return v;
        }
        if (c == 't' || c == 'f') return boolean();
        if (c == 'n') {
            if (s_.compare(p_, 4, "null") != 0) error("expected null");
            p_ += 4;
            return JsonValue{};
        }
        return number();
    }

    JsonValue object() {
        JsonValue v;
        v.kind_ = JsonValue::Kind::Object;
        want('{');
        blank();
        if (at() == '}') {
            ++p_;
            return v;
        }
        for (;;) {
            blank();
            std::string key = string();
            blank();
            want(':');
            blank();
            JsonValue member = value();
            bool replaced = false;
            for (auto& kv : v.members_) {
                if (kv.first == key) {  // a repeated key keeps the last value
                    kv.second = std::move(member);
                    replaced = true;
                    break;
                }
            }
            if (!replaced) v.members_.emplace_back(std::move(key), std::move(member));
            blank();
            if (at() == ',') {
                ++p_;
                continue;
            }
            want('}');
            break;
        }
        v.used_.assign(v.members_.size(), false);
        return v;
    }

    JsonValue array() {
        JsonValue v;
        v.kind_ = JsonValue::Kind::Array;
        want('[');
        blank();
        if (at() == ']') {
            ++p_;
            return v;
        }
        for (;;) {
            blank();
            v.items_.push_back(value());
            blank();
            if (at() == ',') {
                ++p_;
                continue;
            }
            want(']');
            break;
        }
        return v;
    }

    std::string string() {
        want('"');
        std::string out;
        for (;;) {
            if (p_ >= s_.size()) error("unterminated string");
            const char c = s_[p_++];
            if (c == '"') break;
            if (c != '\\') {
                out.push_back(c);
                continue;
            }
            if (p_ >= s_.size()) error("unterminated escape");
            const char e = s_[p_++];
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
                    if (p_ + 4 > s_.size()) error("truncated \\u escape");
                    const unsigned cp =
                        static_cast<unsigned>(std::strtoul(s_.substr(p_, 4).c_str(), nullptr, 16));
                    p_ += 4;
                    if (cp < 0x80) {
                        out.push_back(static_cast<char>(cp));
                    } else if (cp < 0x800) {
                        out.push_back(static_cast<char>(0xC0 | (cp >> 6)));
                        out.push_back(static_cast<char>(0x80 | (cp & 0x3F)));
                    } else {
                        out.push_back(static_cast<char>(0xE0 | (cp >> 12)));
                        out.push_back(static_cast<char>(0x80 | ((cp >> 6) & 0x3F)));
                        out.push_back(static_cast<char>(0x80 | (cp & 0x3F)));
                    }
                    break;
                }
                default: error("unknown escape sequence");
            }
        }
        return out;
    }

    JsonValue boolean() {
        JsonValue v;
        v.kind_ = JsonValue::Kind::Bool;
        if (s_.compare(p_, 4, "true") == 0) {
            v.flag_ = true;
            p_ += 4;
        } else if (s_.compare(p_, 5, "false") == 0) {
            p_ += 5;
        } else {
            error("expected true or false");
        }
        return v;
    }

    JsonValue number() {
        const size_t start = p_;
        while (p_ < s_.size()) {
            const char c = s_[p_];
            if (std::isdigit(static_cast<unsigned char>(c)) || c == '-' || c == '+' || c == '.' ||
                c == 'e' || c == 'E') {
                ++p_;
            } else {
                break;
            }
        }
        if (p_ == start) error("expected a value");
        const std::string tok = s_.substr(start, p_ - start);
        char* end = nullptr;
        const double d = std::strtod(tok.c_str(), &end);
        if (end == tok.c_str()) error("malformed number '" + tok + "'");
        JsonValue v;
        v.kind_ = JsonValue::Kind::Number;
        v.num_ = d;
        return v;
    }

    const std::string& s_;
    size_t p_ = 0;
};

JsonValue JsonValue::parse(const std::string& text) { return JsonReader(text).document(); }

JsonValue JsonValue::parse_file(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open configuration file: " + path);
    std::ostringstream ss;
    ss << in.rdbuf();
    return parse(ss.str());
}

const JsonValue& JsonValue::get(const std::string& key) const {
    if (kind_ != Kind::Object) return null_value();
    for (size_t i = 0; i < members_.size(); ++i) {
        if (members_[i].first == key) {
            used_[i] = true;
            return members_[i].second;
        }
    }
    return null_value();
}

JsonValue& JsonValue::member(const std::string& key) {
    if (kind_ != Kind::Object) {
        *this = JsonValue{};
        kind_ = Kind::Object;
    }
    for (auto& kv : members_) {
        if (kv.first == key) return kv.second;
    }
    members_.emplace_back(key, JsonValue{});
    used_.push_back(false);
    return members_.back().second;
}

void JsonValue::unused_keys(const std::string& prefix, std::vector<std::string>& out) const {
    if (kind_ != Kind::Object) return;
    for (size_t i = 0; i < members_.size(); ++i) {
        const std::string path = prefix.empty() ? members_[i].first : prefix + "." + members_[i].first;
        if (!used_[i]) {
            out.push_back(path);
        } else {
            members_[i].second.unused_keys(path, out);
        }
    }
}

void JsonValue::read(const std::string& key, double& out) const {
    const JsonValue& v = get(key);
    if (v.is_number()) out = v.num_;
}
void JsonValue::read(const std::string& key, int& out) const {
    const JsonValue& v = get(key);
    if (v.is_number()) out = static_cast<int>(v.num_);
}
void JsonValue::read(const std::string& key, unsigned& out) const {
    const JsonValue& v = get(key);
    if (v.is_number()) out = static_cast<unsigned>(v.num_);
}
void JsonValue::read(const std::string& key, unsigned long long& out) const {
    const JsonValue& v = get(key);
    if (v.is_number()) out = static_cast<unsigned long long>(v.num_);
}
void JsonValue::read(const std::string& key, unsigned long& out) const {
    const JsonValue& v = get(key);
    if (v.is_number()) out = static_cast<unsigned long>(v.num_);
}
void JsonValue::read(const std::string& key, bool& out) const {
    const JsonValue& v = get(key);
    if (v.is_bool()) out = v.flag_;
}
void JsonValue::read(const std::string& key, std::string& out) const {
    const JsonValue& v = get(key);
    if (v.is_string()) out = v.str_;
}

}  // namespace asg_host
