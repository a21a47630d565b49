#include "config_doc.h"

#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <stdexcept>

namespace asg_host {

namespace {

enum class Tok : uint8_t { End, Open, Close, OpenList, CloseList, Colon, Comma, Str, Num, True, False, Null };

struct Token {
    Tok tok = Tok::End;
    size_t at = 0;     // byte offset of the token
    std::string text;  // decoded string
    double num = 0.0;
};

[[noreturn]] void bad(size_t at, const std::string& why) {
    throw std::runtime_error("JSON parse error at byte " + std::to_string(at) + ": " + why);
}

// Appends code point cp as UTF-8.
void put_utf8(std::string& out, uint32_t cp) {
    if (cp < 0x80) {
        out += static_cast<char>(cp);
    } else if (cp < 0x800) {
        out += static_cast<char>(0xC0 | (cp >> 6));
        out += static_cast<char>(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
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

int hex_digit(char c) {
    if (c >= '0' && c <= '9') return c - '0';
    if (c >= 'a' && c <= 'f') return c - 'a' + 10;
    if (c >= 'A' && c <= 'F') return c - 'A' + 10;
    return -1;
}

// Splits the text into tokens, dropping blanks and // comments.
class Scanner {
public:
    explicit Scanner(const std::string& s) : s_(s) {}

    Token next() {
        skip_blanks();
        Token t;
        t.at = i_;
        if (i_ == s_.size()) {
            return t;
        }
        const char c = s_[i_];
        switch (c) {
            case '{': ++i_; t.tok = Tok::Open; return t;
            case '}': ++i_; t.tok = Tok::Close; return t;
            case '[': ++i_; t.tok = Tok::OpenList; return t;
            case ']': ++i_; t.tok = Tok::CloseList; return t;
            case ':': ++i_; t.tok = Tok::Colon; return t;
            case ',': ++i_; t.tok = Tok::Comma; return t;
            case '"': t.tok = Tok::Str; t.text = string_body(); return t;
            default: break;
        }
        for (const auto& [word, tok] : {std::pair<const char*, Tok>{"true", Tok::True},
                                        {"false", Tok::False}, {"null", Tok::Null}}) {
            const size_t len = std::char_traits<char>::length(word);
            if (s_.compare(i_, len, word) == 0) {
                i_ += len;
                t.tok = tok;
                return t;
            }
        }
        // A number: the longest run of sign, digit, point and exponent
        // characters, converted by strtod (which reads its longest valid
        // prefix, as the reference's loader does).
        const size_t run = s_.find_first_not_of("+-0123456789.eE", i_);
        const size_t end = run == std::string::npos ? s_.size() : run;
        if (end == i_) {
            bad(i_, "no value starts with '" + std::string(1, c) + "'");
        }
        const std::string digits = s_.substr(i_, end - i_);
        char* stop = nullptr;
        t.num = std::strtod(digits.c_str(), &stop);
        if (stop == digits.c_str()) {
            bad(i_, "'" + digits + "' is not a number");
        }
        t.tok = Tok::Num;
        i_ = end;
        return t;
    }

private:
    void skip_blanks() {
        for (;;) {
            i_ = std::min(s_.size(), s_.find_first_not_of(" \t\r\n", i_));
            if (s_.compare(i_, 2, "//") != 0) {
                return;
            }
            const size_t eol = s_.find('\n', i_);
            i_ = eol == std::string::npos ? s_.size() : eol;
        }
    }

    uint32_t four_hex() {
        if (i_ + 4 > s_.size()) {
            bad(i_, "\\u must be followed by four hex digits");
        }
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) {
            const int d = hex_digit(s_[i_ + k]);
            if (d < 0) {
                bad(i_ + k, "\\u must be followed by four hex digits");
            }
            v = v * 16 + static_cast<uint32_t>(d);
        }
        i_ += 4;
        return v;
    }

    std::string string_body() {
        const size_t open = i_++;
        std::string out;
        for (;;) {
            const size_t q = s_.find_first_of("\"\\", i_);
            if (q == std::string::npos) {
                bad(open, "string is not closed");
            }
            out.append(s_, i_, q - i_);
            i_ = q + 1;
            if (s_[q] == '"') {
                return out;
            }
            if (i_ == s_.size()) {
                bad(q, "string is not closed");
            }
            const char e = s_[i_++];
            switch (e) {
                case '"': case '\\': case '/': out += e; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t cp = four_hex();
                    // a high surrogate followed by \u and a low one: one code point
                    if (cp >= 0xD800 && cp < 0xDC00 && s_.compare(i_, 2, "\\u") == 0) {
                        const size_t save = i_;
                        i_ += 2;
                        const uint32_t lo = four_hex();
                        if (lo >= 0xDC00 && lo < 0xE000) {
                            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                        } else {
                            i_ = save;
                        }
                    }
                    put_utf8(out, cp);
                    break;
                }
                default: bad(i_ - 2, std::string("'\\") + e + "' is not an escape");
            }
        }
    }

    const std::string& s_;
    size_t i_ = 0;
};

}  // namespace

// Builds the value tree of one text into the document's arena: a state
// machine over the token stream with a stack of the open containers. A
// container is attached to its parent when it opens, so closing one only
// pops the stack.
class ConfigParser {
public:
    ConfigParser(ConfigDoc& doc, const std::string& text) : doc_(doc), scan_(text) {}

    uint32_t run() {
        enum class Want { Value, FirstMember, Member, Colon, FirstItem, After };
        Want want = Want::Value;
        std::vector<uint32_t> open;  // arena indices of the open containers
        uint32_t root = 0;
        uint32_t member = 0;         // slot of the object member being read
        Token t;
        bool reuse = false;
        for (;;) {
            if (!reuse) {
                t = scan_.next();
            }
            reuse = false;
            const bool in_object =
                !open.empty() && doc_.values_[open.back()].kind == ConfigDoc::Kind::Object;
            switch (want) {
                case Want::Value: {
                    const uint32_t v = make_value(t);
                    if (open.empty()) {
                        root = v;
                    } else if (in_object) {
                        doc_.slots_[member].value = v;
                    } else {
                        ConfigDoc::Slot item;
                        item.value = v;
                        doc_.slots_.push_back(item);
                        doc_.values_[open.back()].slots.push_back(
                            static_cast<uint32_t>(doc_.slots_.size() - 1));
                    }
                    const ConfigDoc::Kind k = doc_.values_[v].kind;
                    if (k == ConfigDoc::Kind::Object) {
                        open.push_back(v);
                        want = Want::FirstMember;
                    } else if (k == ConfigDoc::Kind::Array) {
                        open.push_back(v);
                        want = Want::FirstItem;
                    } else {
                        want = Want::After;
                    }
                    break;
                }
                case Want::FirstMember:
                    if (t.tok == Tok::Close) {
                        open.pop_back();
                        want = Want::After;
                        break;
                    }
                    [[fallthrough]];
                case Want::Member:
                    if (t.tok != Tok::Str) {
                        bad(t.at, "an object member must start with a quoted name");
                    }
                    member = doc_.member_slot(open.back(), t.text);
                    want = Want::Colon;
                    break;
                case Want::Colon:
                    if (t.tok != Tok::Colon) {
                        bad(t.at, "':' must follow a member name");
                    }
                    want = Want::Value;
                    break;
                case Want::FirstItem:
                    if (t.tok == Tok::CloseList) {
                        open.pop_back();
                        want = Want::After;
                    } else {
                        want = Want::Value;
                        reuse = true;
                    }
                    break;
                case Want::After:
                    if (open.empty()) {
                        if (t.tok != Tok::End) {
                            bad(t.at, "text continues after the document");
                        }
                        return root;
                    }
                    if (t.tok == Tok::Comma) {
                        want = in_object ? Want::Member : Want::Value;
                    } else if (t.tok == (in_object ? Tok::Close : Tok::CloseList)) {
                        open.pop_back();
                    } else if (t.tok == Tok::End) {
                        bad(t.at, "the document ends inside an open object or array");
                    } else {
                        bad(t.at, in_object ? "',' or '}' is expected after an object member"
                                            : "',' or ']' is expected after an array item");
                    }
                    break;
            }
        }
    }

private:
    uint32_t make_value(Token& t) {
        switch (t.tok) {
            case Tok::Open: return doc_.add_value(ConfigDoc::Kind::Object);
            case Tok::OpenList: return doc_.add_value(ConfigDoc::Kind::Array);
            case Tok::Str: {
                const uint32_t v = doc_.add_value(ConfigDoc::Kind::String);
                doc_.values_[v].text = std::move(t.text);
                return v;
            }
            case Tok::Num: {
                const uint32_t v = doc_.add_value(ConfigDoc::Kind::Number);
                doc_.values_[v].num = t.num;
                return v;
            }
            case Tok::True:
            case Tok::False: {
                const uint32_t v = doc_.add_value(ConfigDoc::Kind::Bool);
                doc_.values_[v].flag = t.tok == Tok::True;
                return v;
            }
            case Tok::Null: return doc_.add_value(ConfigDoc::Kind::Null);
            case Tok::End: bad(t.at, "the document ends where a value is expected");
            default: bad(t.at, "a value is expected here");
        }
    }

    ConfigDoc& doc_;
    Scanner scan_;
};

uint32_t ConfigDoc::add_value(Kind kind) {
    values_.emplace_back();
    values_.back().kind = kind;
    return static_cast<uint32_t>(values_.size() - 1);
}

uint32_t ConfigDoc::member_slot(uint32_t object, const std::string& key) {
    for (const uint32_t s : values_[object].slots) {
        if (slots_[s].key == key) {
            return s;  // a repeated key: the later value replaces the earlier
        }
    }
    Slot slot;
    slot.key = key;
    slot.value = add_value(Kind::Null);
    slots_.push_back(std::move(slot));
    const uint32_t id = static_cast<uint32_t>(slots_.size() - 1);
    values_[object].slots.push_back(id);
    return id;
}

uint32_t ConfigDoc::parse_into(const std::string& text) { return ConfigParser(*this, text).run(); }

ConfigDoc ConfigDoc::parse(const std::string& text) {
    ConfigDoc doc;
    doc.root_ = doc.parse_into(text);
    return doc;
}

ConfigDoc ConfigDoc::load(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        throw std::runtime_error("cannot open configuration file: " + path);
    }
    const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    return parse(text);
}

ConfigDoc::Node ConfigDoc::Node::at(const std::string& key) const {
    if (kind() != Kind::Object) {
        return Node();
    }
    for (const uint32_t s : doc_->values_[id_].slots) {
        const Slot& slot = doc_->slots_[s];
        if (slot.key == key) {
            slot.asked = true;
            return Node(doc_, slot.value);
        }
    }
    return Node();
}

void ConfigDoc::assign(const std::string& dotted_path, const std::string& literal) {
    const uint32_t value = parse_into(literal);
    uint32_t node = root_;
    size_t from = 0;
    for (;;) {
        if (values_[node].kind != Kind::Object) {
            values_[node] = Value();
            values_[node].kind = Kind::Object;
        }
        const size_t dot = dotted_path.find('.', from);
        const std::string key =
            dotted_path.substr(from, dot == std::string::npos ? std::string::npos : dot - from);
        const uint32_t slot = member_slot(node, key);
        if (dot == std::string::npos) {
            slots_[slot].value = value;
            return;
        }
        node = slots_[slot].value;
        from = dot + 1;
    }
}

void ConfigDoc::unread_of(uint32_t value, const std::string& prefix,
                          std::vector<std::string>& out) const {
    if (values_[value].kind != Kind::Object) {
        return;
    }
    std::vector<uint32_t> order = values_[value].slots;
    std::sort(order.begin(), order.end(),
              [&](uint32_t a, uint32_t b) { return slots_[a].key < slots_[b].key; });
    for (const uint32_t s : order) {
        const Slot& slot = slots_[s];
        const std::string path = prefix.empty() ? slot.key : prefix + "." + slot.key;
        if (!slot.asked) {
            out.push_back(path);
        } else {
            unread_of(slot.value, path, out);
        }
    }
}

std::vector<std::string> ConfigDoc::unread_paths() const {
    std::vector<std::string> out;
    unread_of(root_, "", out);
    return out;
}

}  // namespace asg_host
