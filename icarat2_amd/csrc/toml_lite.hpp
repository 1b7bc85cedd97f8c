// toml_lite.hpp -- the subset of TOML that ICARAT's configuration files use:
// [table] headers, key = value pairs, comments, and values that are integers,
// floats, booleans, basic strings or (nested, multi-line) arrays.  Inline
// tables, dates and dotted keys are rejected with an error.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace icarat_hip {
namespace toml {

struct Value {
  enum Kind { kInt, kFloat, kBool, kString, kArray } kind = kInt;
  long long i = 0;
  double f = 0.0;
  bool b = false;
  std::string s;
  std::vector<Value> a;

  bool is_number() const { return kind == kInt || kind == kFloat; }
  double number() const;              // int or float, else throws
  const Value &at(size_t k) const;    // array element, else throws
  size_t size() const { return a.size(); }
};

// table name -> key -> value (keys before any header go to table "")
using Document = std::map<std::string, std::map<std::string, Value>>;

// Parses text; throws std::runtime_error("line N: ...") on malformed input.
Document parse(const std::string &text);
Document parse_file(const std::string &path);

// Lookup helpers (nullptr when absent)
const Value *find(const Document &d, const std::string &table, const std::string &key);

}  // namespace toml
}  // namespace icarat_hip
