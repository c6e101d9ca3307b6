#pragma once
// Minimal JSON reader for the .pmesh header (objects, arrays, strings, numbers, literals).

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "core.hpp"

namespace gsk::json {

struct Value {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Value> arr;
  std::map<std::string, Value> obj;

  const Value& at(const std::string& k) const {
    auto it = obj.find(k);
    GSK_CHECK(kind == Object && it != obj.end(), "pmesh header: missing key \"" + k + "\"");
    return it->second;
  }
  bool has(const std::string& k) const { return kind == Object && obj.count(k) != 0; }
  long long as_int() const {
    GSK_CHECK(kind == Number, "pmesh header: expected a number");
    return static_cast<long long>(num);
  }
  const std::string& as_str() const {
    GSK_CHECK(kind == String, "pmesh header: expected a string");
    return str;
  }
};

class Parser {
 public:
  explicit Parser(const std::string& s) : s_(s) {}
  Value parse() {
    Value v = value();
    ws();
    GSK_CHECK(i_ == s_.size(), "pmesh header: trailing characters after JSON");
    return v;
  }

 private:
  void ws() {
    while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\n' || s_[i_] == '\r' || s_[i_] == '\t')) ++i_;
  }
  char peek() {
    ws();
    GSK_CHECK(i_ < s_.size(), "pmesh header: unexpected end of JSON");
    return s_[i_];
  }
  void expect(char c) {
    GSK_CHECK(peek() == c, strf("pmesh header: expected '%c' at %zu", c, i_));
    ++i_;
  }
  std::string string_lit() {
    expect('"');
    std::string out;
    while (true) {
      GSK_CHECK(i_ < s_.size(), "pmesh header: unterminated string");
      const char c = s_[i_++];
      if (c == '"') break;
      if (c == '\\') {
        GSK_CHECK(i_ < s_.size(), "pmesh header: bad escape");
        const char e = s_[i_++];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {
            GSK_CHECK(i_ + 4 <= s_.size(), "pmesh header: bad \\u escape");
            const unsigned cp = static_cast<unsigned>(std::stoul(s_.substr(i_, 4), nullptr, 16));
            i_ += 4;
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
          default: out += e; break;
        }
      } else {
        out += c;
      }
    }
    return out;
  }
  Value value() {
    const char c = peek();
    Value v;
    if (c == '{') {
      ++i_;
      v.kind = Value::Object;
      if (peek() == '}') { ++i_; return v; }
      while (true) {
        std::string k = string_lit();
        expect(':');
        v.obj[k] = value();
        if (peek() == ',') { ++i_; continue; }
        expect('}');
        break;
      }
    } else if (c == '[') {
      ++i_;
      v.kind = Value::Array;
      if (peek() == ']') { ++i_; return v; }
      while (true) {
        v.arr.push_back(value());
        if (peek() == ',') { ++i_; continue; }
        expect(']');
        break;
      }
    } else if (c == '"') {
      v.kind = Value::String;
      v.str = string_lit();
    } else if (s_.compare(i_, 4, "true") == 0) {
      i_ += 4; v.kind = Value::Bool; v.b = true;
    } else if (s_.compare(i_, 5, "false") == 0) {
      i_ += 5; v.kind = Value::Bool; v.b = false;
    } else if (s_.compare(i_, 4, "null") == 0) {
      i_ += 4; v.kind = Value::Null;
    } else {
      std::size_t used = 0;
      v.kind = Value::Number;
      v.num = std::stod(s_.substr(i_, 64), &used);
      GSK_CHECK(used > 0, "pmesh header: bad number");
      i_ += used;
    }
    return v;
  }
  const std::string& s_;
  std::size_t i_ = 0;
};

inline Value parse(const std::string& s) { return Parser(s).parse(); }

}  // namespace gsk::json
