#include "json.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace qiotk {

namespace {

struct Parser {
  const char* p;
  const char* end;

  [[noreturn]] void error(const char* what) const {
    throw JsonError(std::string("JSON parse error: ") + what);
  }
  void ws() {
    while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
  }
  bool lit(const char* s) {
    size_t n = std::strlen(s);
    if ((size_t)(end - p) >= n && std::memcmp(p, s, n) == 0) {
      p += n;
      return true;
    }
    return false;
  }
  static void put_utf8(std::string& out, uint32_t cp) {
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
  uint32_t hex4() {
    if (end - p < 4) error("bad \\u escape");
    uint32_t v = 0;
    for (int k = 0; k < 4; ++k) {
      char c = *p++;
      v <<= 4;
      if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
      else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
      else error("bad \\u escape");
    }
    return v;
  }
  std::string str() {
    if (p >= end || *p != '"') error("expected string");
    ++p;
    std::string out;
    while (true) {
      if (p >= end) error("unterminated string");
      char c = *p++;
      if (c == '"') break;
      if (c == '\\') {
        if (p >= end) error("bad escape");
        char e = *p++;
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
            uint32_t cp = hex4();
            if (cp >= 0xD800 && cp < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
              p += 2;
              uint32_t lo = hex4();
              cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
            }
            put_utf8(out, cp);
            break;
          }
          default: error("bad escape");
        }
      } else {
        out += c;
      }
    }
    return out;
  }
  Json number() {
    const char* s = p;
    bool is_int = true;
    if (p < end && (*p == '-' || *p == '+')) ++p;
    while (p < end && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' ||
                       *p == '-' || *p == '+')) {
      if (*p == '.' || *p == 'e' || *p == 'E') is_int = false;
      ++p;
    }
    std::string tok(s, p);
    if (tok.empty() || tok == "-") error("bad number");
    char* q = nullptr;
    if (is_int && tok.size() < 19) {
      long long v = std::strtoll(tok.c_str(), &q, 10);
      if (*q) error("bad number");
      return Json((int64_t)v);
    }
    double d = std::strtod(tok.c_str(), &q);
    if (*q) error("bad number");
    return Json(d);
  }
  Json value(int depth) {
    if (depth > 512) error("nesting too deep");
    ws();
    if (p >= end) error("unexpected end");
    char c = *p;
    if (c == '{') {
      ++p;
      Json o = Json::object();
      ws();
      if (p < end && *p == '}') {
        ++p;
        return o;
      }
      while (true) {
        ws();
        std::string k = str();
        ws();
        if (p >= end || *p != ':') error("expected ':'");
        ++p;
        o[k] = value(depth + 1);
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == '}') {
          ++p;
          return o;
        }
        error("expected ',' or '}'");
      }
    }
    if (c == '[') {
      ++p;
      Json a = Json::array();
      ws();
      if (p < end && *p == ']') {
        ++p;
        return a;
      }
      while (true) {
        a.push_back(value(depth + 1));
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == ']') {
          ++p;
          return a;
        }
        error("expected ',' or ']'");
      }
    }
    if (c == '"') return Json(str());
    if (lit("true")) return Json(true);
    if (lit("false")) return Json(false);
    if (lit("null")) return Json();
    return number();
  }
};

void dump_string(const std::string& s, std::string& out) {
  out += '"';
  for (unsigned char c : s) {
    switch (c) {
      case '"': out += "\\\""; break;
      case '\\': out += "\\\\"; break;
      case '\n': out += "\\n"; break;
      case '\r': out += "\\r"; break;
      case '\t': out += "\\t"; break;
      default:
        if (c < 0x20) {
          char buf[8];
          std::snprintf(buf, sizeof buf, "\\u%04x", c);
          out += buf;
        } else {
          out += (char)c;
        }
    }
  }
  out += '"';
}

}  // namespace

Json Json::parse(const std::string& text) {
  Parser ps{text.data(), text.data() + text.size()};
  Json v = ps.value(0);
  ps.ws();
  if (ps.p != ps.end) ps.error("trailing characters");
  return v;
}

bool Json::as_bool() const {
  if (type_ != BOOL) throw JsonError("expected boolean");
  return b_;
}
double Json::as_double() const {
  if (type_ != NUMBER) throw JsonError("expected number");
  return is_int_ ? (double)i_ : d_;
}
int64_t Json::as_int() const {
  if (type_ != NUMBER) throw JsonError("expected number");
  if (is_int_) return i_;
  if (d_ != std::floor(d_)) throw JsonError("expected integer");
  return (int64_t)d_;
}
const std::string& Json::as_string() const {
  if (type_ != STRING) throw JsonError("expected string");
  return s_;
}
size_t Json::size() const {
  if (type_ == ARRAY) return arr_.size();
  if (type_ == OBJECT) return obj_.size();
  throw JsonError("expected array");
}
const Json& Json::operator[](size_t i) const {
  if (type_ != ARRAY || i >= arr_.size()) throw JsonError("array index out of range");
  return arr_[i];
}
Json& Json::operator[](size_t i) {
  if (type_ != ARRAY || i >= arr_.size()) throw JsonError("array index out of range");
  return arr_[i];
}
void Json::push_back(Json v) {
  if (type_ == NUL) type_ = ARRAY;
  if (type_ != ARRAY) throw JsonError("expected array");
  arr_.push_back(std::move(v));
}
bool Json::has(const std::string& key) const {
  if (type_ != OBJECT) return false;
  for (auto& kv : obj_)
    if (kv.first == key) return true;
  return false;
}
const Json& Json::at(const std::string& key) const {
  if (type_ == OBJECT)
    for (auto& kv : obj_)
      if (kv.first == key) return kv.second;
  throw JsonError("missing field `" + key + "`");
}
Json& Json::operator[](const std::string& key) {
  if (type_ == NUL) type_ = OBJECT;
  if (type_ != OBJECT) throw JsonError("expected object");
  for (auto& kv : obj_)
    if (kv.first == key) return kv.second;
  obj_.emplace_back(key, Json());
  return obj_.back().second;
}
void Json::append(std::string key, Json v) {
  if (type_ == NUL) type_ = OBJECT;
  if (type_ != OBJECT) throw JsonError("expected object");
  obj_.emplace_back(std::move(key), std::move(v));
}

void Json::remove(const std::string& key) {
  if (type_ != OBJECT) return;
  for (auto it = obj_.begin(); it != obj_.end(); ++it)
    if (it->first == key) {
      obj_.erase(it);
      return;
    }
}

std::string Json::dump() const {
  std::string out;
  dump_to(out);
  return out;
}

void Json::dump_to(std::string& out) const {
  switch (type_) {
    case NUL: out += "null"; break;
    case BOOL: out += b_ ? "true" : "false"; break;
    case NUMBER: {
      char buf[40];
      if (is_int_) {
        std::snprintf(buf, sizeof buf, "%lld", (long long)i_);
      } else if (std::isfinite(d_) && d_ == std::floor(d_) && std::fabs(d_) < 1e15) {
        std::snprintf(buf, sizeof buf, "%.0f", d_);
      } else if (std::isfinite(d_)) {
        std::snprintf(buf, sizeof buf, "%.17g", d_);
      } else {
        std::snprintf(buf, sizeof buf, "null");
      }
      out += buf;
      break;
    }
    case STRING: dump_string(s_, out); break;
    case ARRAY: {
      out += '[';
      for (size_t i = 0; i < arr_.size(); ++i) {
        if (i) out += ", ";
        arr_[i].dump_to(out);
      }
      out += ']';
      break;
    }
    case OBJECT: {
      out += '{';
      for (size_t i = 0; i < obj_.size(); ++i) {
        if (i) out += ", ";
        dump_string(obj_[i].first, out);
        out += ": ";
        obj_[i].second.dump_to(out);
      }
      out += '}';
      break;
    }
  }
}

}  // namespace qiotk
