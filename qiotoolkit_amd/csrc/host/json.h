// Minimal JSON value, reader and writer for the solver/model/params API.
// The reference parses with rapidjson (utils/json.h:9-15); only the subset the
// request and response documents use is needed here: objects (insertion
// order kept), arrays, strings, numbers, booleans, null.
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace qiotk {

class JsonError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

class Json {
 public:
  enum Type { NUL, BOOL, NUMBER, STRING, ARRAY, OBJECT };

  Json() : type_(NUL) {}
  Json(std::nullptr_t) : type_(NUL) {}
  Json(bool b) : type_(BOOL), b_(b) {}
  Json(double d) : type_(NUMBER), d_(d) {}
  Json(int v) : type_(NUMBER), d_(v), is_int_(true), i_(v) {}
  Json(int64_t v) : type_(NUMBER), d_((double)v), is_int_(true), i_(v) {}
  Json(uint32_t v) : type_(NUMBER), d_(v), is_int_(true), i_(v) {}
  Json(uint64_t v) : type_(NUMBER), d_((double)v), is_int_(true), i_((int64_t)v) {}
  Json(const char* s) : type_(STRING), s_(s) {}
  Json(const std::string& s) : type_(STRING), s_(s) {}

  static Json array() {
    Json j;
    j.type_ = ARRAY;
    return j;
  }
  static Json object() {
    Json j;
    j.type_ = OBJECT;
    return j;
  }
  static Json parse(const std::string& text);

  Type type() const { return type_; }
  bool is_null() const { return type_ == NUL; }
  bool is_bool() const { return type_ == BOOL; }
  bool is_number() const { return type_ == NUMBER; }
  bool is_int() const { return type_ == NUMBER && is_int_; }
  bool is_string() const { return type_ == STRING; }
  bool is_array() const { return type_ == ARRAY; }
  bool is_object() const { return type_ == OBJECT; }

  bool as_bool() const;
  double as_double() const;
  int64_t as_int() const;
  const std::string& as_string() const;

  // arrays
  size_t size() const;
  const Json& operator[](size_t i) const;
  Json& operator[](size_t i);
  void push_back(Json v);
  const std::vector<Json>& items() const { return arr_; }

  // objects
  bool has(const std::string& key) const;
  const Json& at(const std::string& key) const;
  Json& operator[](const std::string& key);  // inserts null if absent
  // append a member whose key the caller knows to be new (no lookup)
  void append(std::string key, Json v);
  const Json& operator[](const std::string& key) const { return at(key); }
  void remove(const std::string& key);
  const std::vector<std::pair<std::string, Json>>& members() const { return obj_; }

  std::string dump() const;
  void dump_to(std::string& out) const;

 private:
  Type type_;
  bool b_ = false;
  double d_ = 0.0;
  bool is_int_ = false;
  int64_t i_ = 0;
  std::string s_;
  std::vector<Json> arr_;
  std::vector<std::pair<std::string, Json>> obj_;
};

}  // namespace qiotk
