// json_min.hpp — a small JSON reader for input decks (RFC 8259 values:
// objects, arrays, numbers, strings, true/false/null).  Objects keep their
// keys in a std::map; a missing key reads as null, like the reference's
// readValue on a null nlohmann::json (Input.h:12-19).
#pragma once

#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace wfj {

struct Value {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Value> arr;
  std::map<std::string, Value> obj;

  bool is_null() const { return kind == Null; }
  // object member or null; array element or null
  const Value& operator[](const std::string& k) const {
    static const Value null_value;
    if (kind != Object) return null_value;
    auto it = obj.find(k);
    return it == obj.end() ? null_value : it->second;
  }
  const Value& operator[](size_t i) const {
    static const Value null_value;
    if (kind != Array || i >= arr.size()) return null_value;
    return arr[i];
  }
  size_t size() const { return kind == Array ? arr.size() : kind == Object ? obj.size() : 0; }
};

class Parser {
 public:
  explicit Parser(const std::string& text) : s_(text) {}
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
    size_t line = 1;
    for (size_t i = 0; i < p_ && i < s_.size(); i++) line += s_[i] == '\n';
    throw std::runtime_error(std::string("JSON: ") + what + " at line " + std::to_string(line));
  }
  void ws() {
    while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\n' || s_[p_] == '\r'))
      p_++;
  }
  bool eat(char c) {
    ws();
    if (p_ < s_.size() && s_[p_] == c) {
      p_++;
      return true;
    }
    return false;
  }
  void expect(char c) {
    if (!eat(c)) fail((std::string("expected '") + c + "'").c_str());
  }
  bool word(const char* w) {
    size_t n = std::char_traits<char>::length(w);
    if (s_.compare(p_, n, w) == 0) {
      p_ += n;
      return true;
    }
    return false;
  }
  Value value() {
    ws();
    if (p_ >= s_.size()) fail("unexpected end");
    Value v;
    const char c = s_[p_];
    if (c == '{') {
      p_++;
      v.kind = Value::Object;
      if (eat('}')) return v;
      do {
        ws();
        if (p_ >= s_.size() || s_[p_] != '"') fail("expected a key");
        std::string k = string();
        expect(':');
        v.obj[k] = value();
      } while (eat(','));
      expect('}');
    } else if (c == '[') {
      p_++;
      v.kind = Value::Array;
      if (eat(']')) return v;
      do v.arr.push_back(value());
      while (eat(','));
      expect(']');
    } else if (c == '"') {
      v.kind = Value::String;
      v.str = string();
    } else if (word("true")) {
      v.kind = Value::Bool;
      v.b = true;
    } else if (word("false")) {
      v.kind = Value::Bool;
    } else if (word("null")) {
    } else if (c == '-' || (c >= '0' && c <= '9')) {
      const char* b = s_.c_str() + p_;
      char* e = nullptr;
      v.num = std::strtod(b, &e);
      if (e == b) fail("bad number");
      p_ += (size_t)(e - b);
      v.kind = Value::Number;
    } else {
      fail("unexpected character");
    }
    return v;
  }
  static void utf8(std::string& o, unsigned cp) {
    if (cp < 0x80) {
      o += (char)cp;
    } else if (cp < 0x800) {
      o += (char)(0xC0 | (cp >> 6));
      o += (char)(0x80 | (cp & 0x3F));
    } else {
      o += (char)(0xE0 | (cp >> 12));
      o += (char)(0x80 | ((cp >> 6) & 0x3F));
      o += (char)(0x80 | (cp & 0x3F));
    }
  }
  std::string string() {
    p_++;  // opening quote
    std::string o;
    while (p_ < s_.size() && s_[p_] != '"') {
      char c = s_[p_++];
      if (c != '\\') {
        o += c;
        continue;
      }
      if (p_ >= s_.size()) fail("bad escape");
      char e = s_[p_++];
      switch (e) {
        case '"': o += '"'; break;
        case '\\': o += '\\'; break;
        case '/': o += '/'; break;
        case 'b': o += '\b'; break;
        case 'f': o += '\f'; break;
        case 'n': o += '\n'; break;
        case 'r': o += '\r'; break;
        case 't': o += '\t'; break;
        case 'u': {
          if (p_ + 4 > s_.size()) fail("bad \\u escape");
          utf8(o, (unsigned)std::stoul(s_.substr(p_, 4), nullptr, 16));
          p_ += 4;
          break;
        }
        default: fail("bad escape");
      }
    }
    if (p_ >= s_.size()) fail("unterminated string");
    p_++;
    return o;
  }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

// readValue / readVector / readArray of Input.h:12-78: false when null
inline bool read(const Value& j, double& v) {
  if (j.kind == Value::Number) { v = j.num; return true; }
  if (j.kind == Value::Bool) { v = j.b ? 1.0 : 0.0; return true; }
  return false;
}
inline bool read(const Value& j, int& v) {
  double d;
  if (!read(j, d)) return false;
  v = (int)d;
  return true;
}
inline bool read(const Value& j, bool& v) {
  if (j.kind == Value::Bool) { v = j.b; return true; }
  if (j.kind == Value::Number) { v = j.num != 0.0; return true; }
  return false;
}
inline bool read(const Value& j, std::string& v) {
  if (j.kind != Value::String) return false;
  v = j.str;
  return true;
}
inline bool read3(const Value& j, double v[3]) {
  if (j.kind != Value::Array || j.arr.size() < 3) return false;
  for (int c = 0; c < 3; c++)
    if (!read(j.arr[c], v[c])) return false;
  return true;
}
inline bool read_array(const Value& j, std::vector<double>& v) {
  if (j.kind != Value::Array) return false;
  v.clear();
  for (const Value& e : j.arr) {
    double d = 0.0;
    read(e, d);
    v.push_back(d);
  }
  return true;
}

}  // namespace wfj
