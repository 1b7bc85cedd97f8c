// toml_lite.cpp -- see toml_lite.hpp.
#include "toml_lite.hpp"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace icarat_hip {
namespace toml {

double Value::number() const {
  if (kind == kInt) return (double)i;
  if (kind == kFloat) return f;
  throw std::runtime_error("expected a number");
}

const Value &Value::at(size_t k) const {
  if (kind != kArray || k >= a.size()) throw std::runtime_error("array index out of range");
  return a[k];
}

namespace {

class Parser {
 public:
  explicit Parser(const std::string &t) : t_(t) {}

  Document run() {
    Document doc;
    std::string table;
    doc[table];
    for (;;) {
      skip_blank_lines();
      if (pos_ >= t_.size()) break;
      if (t_[pos_] == '[') {
        ++pos_;
        if (peek() == '[') fail("arrays of tables are not supported");
        skip_ws();
        table = read_key();
        skip_ws();
        expect(']');
        doc[table];
      } else {
        const std::string key = read_key();
        skip_ws();
        expect('=');
        skip_ws();
        Value v = read_value();
        if (doc[table].count(key)) fail("duplicate key '" + key + "'");
        doc[table][key] = std::move(v);
      }
      end_of_line();
    }
    return doc;
  }

 private:
  const std::string &t_;
  size_t pos_ = 0;
  int line_ = 1;

  [[noreturn]] void fail(const std::string &m) const {
    throw std::runtime_error("line " + std::to_string(line_) + ": " + m);
  }
  char peek() const { return pos_ < t_.size() ? t_[pos_] : '\0'; }
  void expect(char c) {
    if (peek() != c) fail(std::string("expected '") + c + "'");
    ++pos_;
  }
  void skip_ws() {
    while (pos_ < t_.size() && (t_[pos_] == ' ' || t_[pos_] == '\t' || t_[pos_] == '\r')) ++pos_;
  }
  void skip_comment() {
    if (peek() == '#')
      while (pos_ < t_.size() && t_[pos_] != '\n') ++pos_;
  }
  // whitespace, comments and newlines (inside arrays and between statements)
  void skip_blank_lines() {
    for (;;) {
      skip_ws();
      skip_comment();
      if (peek() == '\n') {
        ++pos_;
        ++line_;
        continue;
      }
      break;
    }
  }
  void end_of_line() {
    skip_ws();
    skip_comment();
    if (pos_ < t_.size()) {
      if (t_[pos_] != '\n') fail("unexpected text after a value");
      ++pos_;
      ++line_;
    }
  }
  std::string read_key() {
    if (peek() == '"') return read_string();
    size_t b = pos_;
    while (pos_ < t_.size() && (std::isalnum((unsigned char)t_[pos_]) || t_[pos_] == '_' ||
                                t_[pos_] == '-'))
      ++pos_;
    if (pos_ == b) fail("expected a key");
    if (peek() == '.') fail("dotted keys are not supported");
    return t_.substr(b, pos_ - b);
  }
  std::string read_string() {
    expect('"');
    std::string out;
    while (pos_ < t_.size() && t_[pos_] != '"') {
      char c = t_[pos_++];
      if (c == '\n') fail("newline in a string");
      if (c == '\\') {
        const char e = peek();
        ++pos_;
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case '"': out += '"'; break;
          case '\\': out += '\\'; break;
          default: fail("unsupported escape in a string");
        }
      } else {
        out += c;
      }
    }
    expect('"');
    return out;
  }
  Value read_value() {
    Value v;
    const char c = peek();
    if (c == '[') {
      ++pos_;
      v.kind = Value::kArray;
      for (;;) {
        skip_blank_lines();
        if (peek() == ']') {
          ++pos_;
          break;
        }
        v.a.push_back(read_value());
        skip_blank_lines();
        if (peek() == ',') {
          ++pos_;
          continue;
        }
        if (peek() == ']') {
          ++pos_;
          break;
        }
        fail("expected ',' or ']' in an array");
      }
      return v;
    }
    if (c == '"') {
      v.kind = Value::kString;
      v.s = read_string();
      return v;
    }
    if (c == '{') fail("inline tables are not supported");
    if (t_.compare(pos_, 4, "true") == 0) {
      pos_ += 4;
      v.kind = Value::kBool;
      v.b = true;
      return v;
    }
    if (t_.compare(pos_, 5, "false") == 0) {
      pos_ += 5;
      v.kind = Value::kBool;
      v.b = false;
      return v;
    }
    // number: [+-] digits [. digits] [eE [+-] digits], underscores allowed
    size_t b = pos_;
    std::string num;
    bool isfloat = false;
    while (pos_ < t_.size()) {
      const char d = t_[pos_];
      if (std::isdigit((unsigned char)d) || d == '+' || d == '-') {
        num += d;
      } else if (d == '.' || d == 'e' || d == 'E') {
        num += d;
        isfloat = true;
      } else if (d == '_') {
      } else {
        break;
      }
      ++pos_;
    }
    if (pos_ == b || num.empty()) fail("expected a value");
    char *end = nullptr;
    if (isfloat) {
      v.kind = Value::kFloat;
      v.f = std::strtod(num.c_str(), &end);
    } else {
      v.kind = Value::kInt;
      v.i = std::strtoll(num.c_str(), &end, 10);
    }
    if (!end || *end != '\0') fail("malformed number '" + num + "'");
    return v;
  }
};

}  // namespace

Document parse(const std::string &text) { return Parser(text).run(); }

Document parse_file(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open '" + path + "'");
  std::stringstream ss;
  ss << f.rdbuf();
  return parse(ss.str());
}

const Value *find(const Document &d, const std::string &table, const std::string &key) {
  auto t = d.find(table);
  if (t == d.end()) return nullptr;
  auto k = t->second.find(key);
  return k == t->second.end() ? nullptr : &k->second;
}

}  // namespace toml
}  // namespace icarat_hip
