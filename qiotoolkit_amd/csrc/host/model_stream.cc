// Streaming reader of cost-function documents: one pass over the text (or
// over a file in 64 KiB chunks, as the reference's rapidjson FileReadStream
// reads it, utils/stream_handler_json.h:124-182) that writes the terms
// straight into a term CSR.  The tree parser (Json::parse + parse_model)
// stays the definition of the format: this reader only takes documents of
// the usual shape and hands anything else back to it (std::nullopt), so
// errors and their messages are the tree parser's.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

#include "model.h"
#include "proto.h"

namespace qiotk {

namespace {

// Bytes of the document, from memory or refilled from a file.
class Source {
 public:
  Source(const char* text, size_t n) : p_(text), e_(text + n) {}
  explicit Source(FILE* f) : f_(f), buf_(65536) { refill(); }
  int peek() {
    if (p_ == e_ && !refill()) return -1;
    return (unsigned char)*p_;
  }
  int get() {
    const int c = peek();
    if (c >= 0) ++p_;
    return c;
  }
  void ws() {
    while (true) {
      const int c = peek();
      if (c == ' ' || c == '\n' || c == '\r' || c == '\t') ++p_;
      else return;
    }
  }
  bool read_error() const { return f_ && std::ferror(f_); }

 private:
  bool refill() {
    if (!f_) return false;
    const size_t n = std::fread(buf_.data(), 1, buf_.size(), f_);
    p_ = buf_.data();
    e_ = p_ + n;
    return n > 0;
  }
  FILE* f_ = nullptr;
  std::vector<char> buf_;
  const char* p_ = nullptr;
  const char* e_ = nullptr;
};

// Thrown when the document leaves the shape this reader takes; the caller
// falls back to the tree parser.
struct NotStreamable {};

struct Number {
  bool is_int = false;
  int64_t i = 0;
  double d = 0.0;
  double as_double() const { return is_int ? (double)i : d; }
  // Json::as_int: integers, or doubles without a fraction
  int64_t as_int() const {
    if (is_int) return i;
    if (d != (double)(int64_t)d) throw NotStreamable{};
    return (int64_t)d;
  }
};

class Reader {
 public:
  explicit Reader(Source& s) : s_(s) {}

  void expect(int ch) {
    s_.ws();
    if (s_.get() != ch) throw NotStreamable{};
  }
  // the token Json::parse reads as a number, converted the same way
  Number number() {
    s_.ws();
    char tok[64];
    size_t n = 0;
    bool is_int = true;
    int c = s_.peek();
    if (c == '-' || c == '+') tok[n++] = (char)s_.get();
    while (true) {
      c = s_.peek();
      if (!((c >= '0' && c <= '9') || c == '.' || c == 'e' || c == 'E' || c == '-' || c == '+')) break;
      if (c == '.' || c == 'e' || c == 'E') is_int = false;
      if (n + 1 >= sizeof tok) throw NotStreamable{};
      tok[n++] = (char)s_.get();
    }
    tok[n] = 0;
    if (n == 0 || (n == 1 && tok[0] == '-')) throw NotStreamable{};
    Number v;
    char* q = nullptr;
    if (is_int && n < 19) {
      v.is_int = true;
      v.i = std::strtoll(tok, &q, 10);
    } else {
      v.d = std::strtod(tok, &q);
    }
    if (*q) throw NotStreamable{};
    return v;
  }
  // strings without escapes (keys, type, version, ids of the initial
  // configuration); an escape hands the document to the tree parser
  std::string string() {
    s_.ws();
    if (s_.get() != '"') throw NotStreamable{};
    std::string out;
    while (true) {
      const int c = s_.get();
      if (c < 0 || c == '\\') throw NotStreamable{};
      if (c == '"') return out;
      out += (char)c;
    }
  }
  // any value, skipped (escapes in skipped strings are stepped over)
  void skip(int depth = 0) {
    if (depth > 512) throw NotStreamable{};
    s_.ws();
    const int c = s_.peek();
    if (c == '{' || c == '[') {
      const int close = c == '{' ? '}' : ']';
      s_.get();
      s_.ws();
      if (s_.peek() == close) {
        s_.get();
        return;
      }
      while (true) {
        if (c == '{') {
          skip_string();
          expect(':');
        }
        skip(depth + 1);
        s_.ws();
        const int d = s_.get();
        if (d == close) return;
        if (d != ',') throw NotStreamable{};
      }
    }
    if (c == '"') {
      skip_string();
      return;
    }
    if (c == 't' || c == 'f' || c == 'n') {
      const char* lit = c == 't' ? "true" : c == 'f' ? "false" : "null";
      for (const char* q = lit; *q; ++q)
        if (s_.get() != *q) throw NotStreamable{};
      return;
    }
    number();
  }
  // ',' (true) or the closing bracket (false)
  bool more(int close) {
    s_.ws();
    const int c = s_.get();
    if (c == ',') return true;
    if (c == close) return false;
    throw NotStreamable{};
  }
  bool empty(int close) {
    s_.ws();
    if (s_.peek() == close) {
      s_.get();
      return true;
    }
    return false;
  }
  void end() {
    s_.ws();
    if (s_.peek() >= 0 || s_.read_error()) throw NotStreamable{};
  }

 private:
  void skip_string() {
    s_.ws();
    if (s_.get() != '"') throw NotStreamable{};
    while (true) {
      const int c = s_.get();
      if (c < 0) throw NotStreamable{};
      if (c == '"') return;
      if (c == '\\' && s_.get() < 0) throw NotStreamable{};
    }
  }
  Source& s_;
};

// {"cost_function": {...}, other keys...}
Model read_document(Source& src) {
  Reader r(src);
  std::string type, version = "1.0";
  bool seen_cf = false, seen_type = false, seen_version = false, seen_terms = false, seen_ic = false;
  std::vector<int64_t> off(1, 0);
  std::vector<int32_t> ids;
  std::vector<double> coef;
  std::vector<std::pair<std::string, int64_t>> ic;
  r.expect('{');
  if (r.empty('}')) throw NotStreamable{};
  do {
    const std::string key = r.string();
    r.expect(':');
    if (key == "type" || key == "version" || key == "terms" || key == "initial_configuration")
      throw NotStreamable{};  // the cost function at the top level: the tree parser
    if (key != "cost_function") {
      r.skip();
      continue;
    }
    if (seen_cf) throw NotStreamable{};
    seen_cf = true;
    r.expect('{');
    if (r.empty('}')) throw NotStreamable{};
    do {
      const std::string k = r.string();
      r.expect(':');
      if (k == "type") {
        if (seen_type) throw NotStreamable{};
        seen_type = true;
        type = r.string();
      } else if (k == "version") {
        if (seen_version) throw NotStreamable{};
        seen_version = true;
        version = r.string();
      } else if (k == "terms") {
        if (seen_terms) throw NotStreamable{};
        seen_terms = true;
        r.expect('[');
        if (r.empty(']')) throw NotStreamable{};  // no terms: the tree parser's error
        do {
          // {"c": number, "ids": [integers]} in either order; other keys skipped
          bool have_c = false, have_ids = false;
          double c = 0.0;
          r.expect('{');
          if (r.empty('}')) throw NotStreamable{};
          do {
            const std::string tk = r.string();
            r.expect(':');
            if (tk == "c") {
              if (have_c) throw NotStreamable{};
              have_c = true;
              c = r.number().as_double();
            } else if (tk == "ids") {
              if (have_ids) throw NotStreamable{};
              have_ids = true;
              r.expect('[');
              if (!r.empty(']')) {
                do {
                  ids.push_back((int32_t)r.number().as_int());
                } while (r.more(']'));
              }
            } else {
              r.skip();
            }
          } while (r.more('}'));
          if (!have_c || !have_ids) throw NotStreamable{};
          coef.push_back(c);
          off.push_back((int64_t)ids.size());
        } while (r.more(']'));
      } else if (k == "initial_configuration") {
        if (seen_ic) throw NotStreamable{};
        seen_ic = true;
        std::unordered_set<std::string> names;
        r.expect('{');
        if (!r.empty('}')) {
          do {
            std::string name = r.string();
            r.expect(':');
            const int64_t v = r.number().as_int();
            if (!names.insert(name).second) throw NotStreamable{};  // duplicate: the last wins in the tree
            ic.emplace_back(std::move(name), v);
          } while (r.more('}'));
        }
      } else {
        r.skip();
      }
    } while (r.more('}'));
  } while (r.more('}'));
  r.end();
  if (!seen_cf || !seen_type || !seen_terms || coef.empty()) throw NotStreamable{};
  int kind;
  if (type == "ising") kind = QMC_ISING;
  else if (type == "pubo") kind = QMC_PUBO;
  else throw NotStreamable{};
  // AbstractIsing/AbstractPubo::match_version (model/ising.h:80-89, model/pubo.h:69-78)
  if (version != "1.0" && version != "1.1") throw NotStreamable{};
  Model m = make_model_flat(kind, (int64_t)coef.size(), off.data(), ids.data(), coef.data());
  m.version = version;
  m.streamed = true;
  if (seen_ic) {
    for (const auto& kv : ic) apply_initial_value(m, kv.first, kv.second);
    finish_initial_configuration(m);
  }
  return m;
}

}  // namespace

std::optional<Model> stream_model_text(const char* text, size_t n) {
  try {
    Source src(text, n);
    return read_document(src);
  } catch (const NotStreamable&) {
    return std::nullopt;
  }
}

std::optional<Model> stream_model_file(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return std::nullopt;
  try {
    Source src(f);
    Model m = read_document(src);
    std::fclose(f);
    return m;
  } catch (const NotStreamable&) {
    std::fclose(f);
    return std::nullopt;
  } catch (...) {
    std::fclose(f);
    throw;
  }
}

Model read_model_text(const char* text) {
  if (auto m = stream_model_text(text, std::strlen(text))) return std::move(*m);
  return parse_model(Json::parse(text));
}

Model read_model_file(const std::string& path) {
  // a protobuf folder (or an unreadable path) fails here and takes the
  // folder reader or the reference's file error
  if (auto m = stream_model_file(path)) return std::move(*m);
  return parse_model(read_input_file(path));
}

}  // namespace qiotk
