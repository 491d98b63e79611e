// json_min.h -- a small JSON reader for the model export's config.json.
//
// Objects, arrays, strings (\" \\ \/ \b \f \n \r \t and \uXXXX below 0x80),
// numbers, true / false / null.  Errors throw std::runtime_error with the
// byte offset.
#pragma once
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace fvc {
namespace json {

struct Value {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Value> arr;
  std::map<std::string, Value> obj;

  bool has(const std::string& k) const { return kind == OBJ && obj.count(k) != 0; }
  const Value& at(const std::string& k) const {
    auto it = obj.find(k);
    if (kind != OBJ || it == obj.end()) throw std::runtime_error("config: missing key '" + k + "'");
    return it->second;
  }
  const Value& at(size_t i) const {
    if (kind != ARR || i >= arr.size()) throw std::runtime_error("config: array index out of range");
    return arr[i];
  }
  double number() const {
    if (kind != NUM) throw std::runtime_error("config: number expected");
    return num;
  }
  bool boolean() const {
    if (kind == BOOL) return b;
    if (kind == NUM) return num != 0.0;
    throw std::runtime_error("config: boolean expected");
  }
  const std::string& string() const {
    if (kind != STR) throw std::runtime_error("config: string expected");
    return str;
  }
  double number_or(const std::string& k, double d) const { return has(k) ? at(k).number() : d; }
};

class Parser {
 public:
  explicit Parser(const std::string& s) : s_(s) {}
  Value parse() {
    Value v = value();
    ws();
    if (p_ != s_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string& s_;
  size_t p_ = 0;

  [[noreturn]] void fail(const char* what) const {
    throw std::runtime_error(std::string("config.json: ") + what + " at byte " + std::to_string(p_));
  }
  void ws() {
    while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\n' || s_[p_] == '\r')) p_++;
  }
  char peek() {
    ws();
    if (p_ >= s_.size()) fail("unexpected end");
    return s_[p_];
  }
  void expect(char c) {
    if (peek() != c) fail("unexpected character");
    p_++;
  }
  bool literal(const char* w) {
    const size_t n = std::char_traits<char>::length(w);
    if (s_.compare(p_, n, w) == 0) {
      p_ += n;
      return true;
    }
    return false;
  }
  Value value() {
    const char c = peek();
    Value v;
    if (c == '{') {
      v.kind = Value::OBJ;
      p_++;
      if (peek() == '}') { p_++; return v; }
      while (true) {
        if (peek() != '"') fail("key expected");
        const std::string k = string_lit();
        expect(':');
        v.obj[k] = value();
        const char d = peek();
        p_++;
        if (d == '}') break;
        if (d != ',') fail("',' or '}' expected");
      }
    } else if (c == '[') {
      v.kind = Value::ARR;
      p_++;
      if (peek() == ']') { p_++; return v; }
      while (true) {
        v.arr.push_back(value());
        const char d = peek();
        p_++;
        if (d == ']') break;
        if (d != ',') fail("',' or ']' expected");
      }
    } else if (c == '"') {
      v.kind = Value::STR;
      v.str = string_lit();
    } else if (literal("true")) {
      v.kind = Value::BOOL;
      v.b = true;
    } else if (literal("false")) {
      v.kind = Value::BOOL;
    } else if (literal("null")) {
      v.kind = Value::NUL;
    } else {
      const char* b = s_.c_str() + p_;
      char* e = nullptr;
      v.num = std::strtod(b, &e);
      if (e == b) fail("value expected");
      v.kind = Value::NUM;
      p_ += (size_t)(e - b);
    }
    return v;
  }
  std::string string_lit() {
    expect('"');
    std::string out;
    while (true) {
      if (p_ >= s_.size()) fail("unterminated string");
      const char c = s_[p_++];
      if (c == '"') break;
      if (c != '\\') { out.push_back(c); continue; }
      if (p_ >= s_.size()) fail("bad escape");
      const char e = s_[p_++];
      switch (e) {
        case '"': case '\\': case '/': out.push_back(e); break;
        case 'b': out.push_back('\b'); break;
        case 'f': out.push_back('\f'); break;
        case 'n': out.push_back('\n'); break;
        case 'r': out.push_back('\r'); break;
        case 't': out.push_back('\t'); break;
        case 'u': {
          if (p_ + 4 > s_.size()) fail("bad \\u escape");
          const long cp = std::strtol(s_.substr(p_, 4).c_str(), nullptr, 16);
          p_ += 4;
          if (cp >= 0x80) fail("non-ASCII \\u escape");
          out.push_back((char)cp);
          break;
        }
        default: fail("bad escape");
      }
    }
    return out;
  }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

}  // namespace json
}  // namespace fvc
