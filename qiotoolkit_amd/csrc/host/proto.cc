// Protobuf problem folders without a protobuf library: the QuantumUtil.Problem
// message (utils/problem.proto:1-30) is read straight from its wire format
// (varints, little-endian doubles, length-delimited fields).
#include "proto.h"

#include <sys/stat.h>

#include <cstdint>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>
#include <vector>

#include "errors.h"

namespace qiotk {

namespace {

// wire types
enum : int { kVarint = 0, kFixed64 = 1, kLen = 2, kFixed32 = 5 };

struct Wire {
  const uint8_t* p;
  const uint8_t* end;

  bool more() const { return p < end; }

  uint64_t varint() {
    uint64_t x = 0;
    for (int shift = 0; shift < 64; shift += 7) {
      if (p >= end) throw Error(kParsingError, "truncated protobuf message");
      const uint8_t c = *p++;
      x |= (uint64_t)(c & 0x7f) << shift;
      if (!(c & 0x80)) return x;
    }
    throw Error(kParsingError, "malformed protobuf varint");
  }
  // next field: number and wire type
  void key(uint32_t& field, int& wire) {
    const uint64_t k = varint();
    field = (uint32_t)(k >> 3);
    wire = (int)(k & 7);
  }
  Wire sub() {
    const uint64_t n = varint();
    if (n > (uint64_t)(end - p)) throw Error(kParsingError, "truncated protobuf message");
    Wire w{p, p + n};
    p += n;
    return w;
  }
  double fixed64_double() {
    if (end - p < 8) throw Error(kParsingError, "truncated protobuf message");
    uint64_t u = 0;
    for (int i = 7; i >= 0; --i) u = (u << 8) | p[i];
    p += 8;
    double d;
    std::memcpy(&d, &u, 8);
    return d;
  }
  void skip(int wire) {
    switch (wire) {
      case kVarint: varint(); return;
      case kFixed64:
        if (end - p < 8) throw Error(kParsingError, "truncated protobuf message");
        p += 8;
        return;
      case kLen: sub(); return;
      case kFixed32:
        if (end - p < 4) throw Error(kParsingError, "truncated protobuf message");
        p += 4;
        return;
      default: throw Error(kParsingError, "unsupported protobuf wire type");
    }
  }
};

// Problem.Term {double c = 1; repeated int64 ids = 2;} (packed or not)
Json read_term(Wire w) {
  double c = 0.0;
  Json ids = Json::array();
  while (w.more()) {
    uint32_t f;
    int wt;
    w.key(f, wt);
    if (f == 1 && wt == kFixed64) {
      c = w.fixed64_double();
    } else if (f == 2 && wt == kLen) {
      Wire packed = w.sub();
      while (packed.more()) ids.push_back(Json((int64_t)packed.varint()));
    } else if (f == 2 && wt == kVarint) {
      ids.push_back(Json((int64_t)w.varint()));
    } else {
      w.skip(wt);
    }
  }
  Json t = Json::object();
  t["c"] = c;
  t["ids"] = ids;
  return t;
}

struct CostFunction {
  bool present = false;
  int64_t type = 0;  // ProblemType: PUBO 0, ISING 1, MAXSAT 2, SOFTSPIN 3
  std::string version;
  std::vector<Json> terms;
};

// Problem {CostFunction cost_function = 1; Struct metadata = 2;};
// CostFunction {ProblemType type = 1; string version = 2; repeated Term terms = 3;
//               map<string, int64> init_config = 4;}
CostFunction read_problem(const std::vector<uint8_t>& bytes) {
  CostFunction cf;
  Wire w{bytes.data(), bytes.data() + bytes.size()};
  while (w.more()) {
    uint32_t f;
    int wt;
    w.key(f, wt);
    if (f != 1 || wt != kLen) {
      w.skip(wt);
      continue;
    }
    cf.present = true;
    Wire c = w.sub();
    while (c.more()) {
      uint32_t g;
      int ct;
      c.key(g, ct);
      if (g == 1 && ct == kVarint) {
        cf.type = (int64_t)c.varint();
      } else if (g == 2 && ct == kLen) {
        Wire s = c.sub();
        cf.version.assign(reinterpret_cast<const char*>(s.p), (size_t)(s.end - s.p));
      } else if (g == 3 && ct == kLen) {
        cf.terms.push_back(read_term(c.sub()));
      } else {
        c.skip(ct);  // init_config is not read by the reference either
      }
    }
  }
  return cf;
}

bool read_file(const std::string& name, std::vector<uint8_t>& out) {
  std::ifstream in(name, std::ios::in | std::ios::binary);
  if (!in) return false;
  out.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
  return true;
}

bool is_folder(const std::string& path) {
  struct stat st;
  return stat(path.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

}  // namespace

Json read_proto_folder(const std::string& path) {
  // the folder name as ProtoReader::Parse derives it (utils/proto_reader.h:45-56)
  const size_t slash = path.find_last_of("/\\");
  const std::string name = (slash != std::string::npos && slash != 0) ? path.substr(slash + 1) : path;
  static const char* kTypes[] = {"pubo", "ising", "maxsat", "softspin"};
  Json terms = Json::array();
  std::string type, version;
  std::vector<uint8_t> bytes;
  for (int k = 0;; ++k) {
    const std::string file = path + "/" + name + "_" + std::to_string(k) + ".pb";
    if (!read_file(file, bytes)) {
      if (k == 0) throw Error(kFileReadError, "Could not open file: " + file);
      break;
    }
    const CostFunction cf = read_problem(bytes);
    if (!cf.present) throw Error(kMissingInput, "Invalid problem message. No cost function found");
    if (k == 0) {
      if (cf.type < 0 || cf.type > 3)
        throw Error(kValueError,
                    "Expected type to be ising, pubo, maxsat or softspin. Invalid problem type specified. ");
      type = kTypes[cf.type];
      version = cf.version;
      if (type == "softspin") {
        if (version != "0.1" && !version.empty())
          throw Error(kValueError, "Expected version to be 0.1 for soft spin. Provided: " + version);
      } else if (version != "1.0" && version != "1.1" && !version.empty()) {
        throw Error(kValueError, "Expected version to be 1.0 or 1.1. Provided: " + version);
      }
      if ((type == "ising" || type == "pubo") && cf.terms.empty())
        throw Error(kValueError, "Problem terms cannot be 0. Please add problem terms");
    }
    for (const Json& t : cf.terms) terms.push_back(t);
  }
  Json c = Json::object();
  c["type"] = type;
  if (!version.empty()) c["version"] = version;
  c["terms"] = terms;
  Json doc = Json::object();
  doc["cost_function"] = c;
  return doc;
}

Json read_input_file(const std::string& path) {
  if (is_folder(path)) return read_proto_folder(path);
  std::ifstream in(path);
  if (!in) throw Error(kFileReadError, "Could not open file: " + path);
  std::stringstream ss;
  ss << in.rdbuf();
  return Json::parse(ss.str());
}

}  // namespace qiotk
