// A small JSON reader for the analysis input deck (RFC 8259 values: null,
// booleans, numbers, strings with escapes, arrays, objects). Objects keep
// their keys in input order. Only what the deck needs: parse, type checks,
// member and element access with errors that name the missing key.
#pragma once

#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace mpm {

class Json {
 public:
  enum Type { Null, Bool, Number, String, Array, Object };

  Json() = default;
  static Json parse(const std::string& text);

  Type type() const { return type_; }
  bool is_null() const { return type_ == Null; }
  bool is_boolean() const { return type_ == Bool; }
  bool is_number() const { return type_ == Number; }
  bool is_string() const { return type_ == String; }
  bool is_array() const { return type_ == Array; }
  bool is_object() const { return type_ == Object; }

  bool contains(const std::string& key) const {
    return type_ == Object && index_.count(key) > 0;
  }
  const Json& at(const std::string& key) const {
    if (type_ != Object) throw std::runtime_error("JSON: not an object (looking up \"" + key + "\")");
    auto it = index_.find(key);
    if (it == index_.end()) throw std::runtime_error("JSON: key \"" + key + "\" not found");
    return items_[it->second].second;
  }
  const Json& at(size_t i) const {
    if (type_ != Array) throw std::runtime_error("JSON: not an array");
    if (i >= elems_.size()) throw std::runtime_error("JSON: index out of range");
    return elems_[i];
  }
  size_t size() const {
    return type_ == Array ? elems_.size() : (type_ == Object ? items_.size() : 0);
  }
  const std::vector<Json>& elements() const { return elems_; }
  const std::vector<std::pair<std::string, Json>>& items() const { return items_; }

  double number() const {
    if (type_ != Number) throw std::runtime_error("JSON: not a number");
    return num_;
  }
  int64_t integer() const {
    const double v = number();
    if (v != std::floor(v)) throw std::runtime_error("JSON: not an integer");
    return (int64_t)v;
  }
  bool boolean() const {
    if (type_ != Bool) throw std::runtime_error("JSON: not a boolean");
    return b_;
  }
  const std::string& str() const {
    if (type_ != String) throw std::runtime_error("JSON: not a string");
    return s_;
  }
  // value or default when the key is absent
  double get(const std::string& key, double dflt) const {
    return contains(key) ? at(key).number() : dflt;
  }

 private:
  friend class JsonParser;
  Type type_ = Null;
  bool b_ = false;
  double num_ = 0.;
  std::string s_;
  std::vector<Json> elems_;
  std::vector<std::pair<std::string, Json>> items_;
  std::map<std::string, size_t> index_;
};

class JsonParser {
 public:
  explicit JsonParser(const std::string& t) : t_(t) {}

  Json value() {
    ws();
    if (i_ >= t_.size()) fail("unexpected end");
    const char c = t_[i_];
    Json j;
    if (c == '{') {
      j.type_ = Json::Object;
      ++i_;
      ws();
      if (peek('}')) return ++i_, j;
      for (;;) {
        ws();
        std::string key = string_lit();
        ws();
        expect(':');
        Json v = value();
        j.index_[key] = j.items_.size();
        j.items_.emplace_back(std::move(key), std::move(v));
        ws();
        if (peek(',')) { ++i_; continue; }
        expect('}');
        return j;
      }
    }
    if (c == '[') {
      j.type_ = Json::Array;
      ++i_;
      ws();
      if (peek(']')) return ++i_, j;
      for (;;) {
        j.elems_.push_back(value());
        ws();
        if (peek(',')) { ++i_; continue; }
        expect(']');
        return j;
      }
    }
    if (c == '"') {
      j.type_ = Json::String;
      j.s_ = string_lit();
      return j;
    }
    if (literal("true")) { j.type_ = Json::Bool; j.b_ = true; return j; }
    if (literal("false")) { j.type_ = Json::Bool; j.b_ = false; return j; }
    if (literal("null")) return j;
    j.type_ = Json::Number;
    j.num_ = number();
    return j;
  }

  void finish() {
    ws();
    if (i_ != t_.size()) fail("trailing characters");
  }

 private:
  const std::string& t_;
  size_t i_ = 0;

  [[noreturn]] void fail(const std::string& what) const {
    throw std::runtime_error("JSON parse error at offset " + std::to_string(i_) + ": " + what);
  }
  void ws() {
    while (i_ < t_.size() && (t_[i_] == ' ' || t_[i_] == '\t' || t_[i_] == '\n' || t_[i_] == '\r'))
      ++i_;
  }
  bool peek(char c) const { return i_ < t_.size() && t_[i_] == c; }
  void expect(char c) {
    if (!peek(c)) fail(std::string("expected '") + c + "'");
    ++i_;
  }
  bool literal(const char* w) {
    const std::string s(w);
    if (t_.compare(i_, s.size(), s) == 0) {
      i_ += s.size();
      return true;
    }
    return false;
  }
  double number() {
    const size_t b = i_;
    if (peek('-')) ++i_;
    while (i_ < t_.size() && (isdigit((unsigned char)t_[i_]) || t_[i_] == '.' || t_[i_] == 'e' ||
                              t_[i_] == 'E' || t_[i_] == '+' || t_[i_] == '-'))
      ++i_;
    if (b == i_) fail("expected a value");
    size_t used = 0;
    const std::string tok = t_.substr(b, i_ - b);
    double v = 0.;
    try {
      v = std::stod(tok, &used);
    } catch (...) {
      fail("bad number '" + tok + "'");
    }
    if (used != tok.size()) fail("bad number '" + tok + "'");
    return v;
  }
  static void utf8(std::string& out, unsigned cp) {
    if (cp < 0x80) {
      out += (char)cp;
    } else if (cp < 0x800) {
      out += (char)(0xC0 | (cp >> 6));
      out += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      out += (char)(0xE0 | (cp >> 12));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    } else {
      out += (char)(0xF0 | (cp >> 18));
      out += (char)(0x80 | ((cp >> 12) & 0x3F));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    }
  }
  unsigned hex4() {
    if (i_ + 4 > t_.size()) fail("short \\u escape");
    unsigned v = 0;
    for (int k = 0; k < 4; ++k) {
      const char c = t_[i_++];
      v <<= 4;
      if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
      else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
      else fail("bad \\u escape");
    }
    return v;
  }
  std::string string_lit() {
    expect('"');
    std::string out;
    for (;;) {
      if (i_ >= t_.size()) fail("unterminated string");
      const char c = t_[i_++];
      if (c == '"') return out;
      if (c != '\\') {
        out += c;
        continue;
      }
      if (i_ >= t_.size()) fail("unterminated escape");
      const char e = t_[i_++];
      switch (e) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          unsigned cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00 && i_ + 1 < t_.size() && t_[i_] == '\\' &&
              t_[i_ + 1] == 'u') {
            i_ += 2;
            const unsigned lo = hex4();
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          utf8(out, cp);
          break;
        }
        default: fail("bad escape");
      }
    }
  }
};

inline Json Json::parse(const std::string& text) {
  JsonParser p(text);
  Json j = p.value();
  p.finish();
  return j;
}

}  // namespace mpm
