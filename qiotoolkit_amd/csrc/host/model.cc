#include "model.h"

#include <algorithm>
#include <cmath>
#include <iterator>
#include <limits>
#include <set>

namespace qiotk {

Model make_model_flat(int kind, int64_t n_terms, const int64_t* offsets, const int32_t* ids,
                      const double* coeffs) {
  // graph::normalize_edges (graph/graph.h:74-177)
  if (n_terms <= 0)
    throw Error(kValueError, "parameter `terms`: size must be greater than 0, found 0");
  Model m;
  m.kind = kind;
  m.term_off.reserve((size_t)n_terms + 1);
  m.term_off.push_back(0);
  m.term_var.reserve((size_t)offsets[n_terms]);
  m.coeff.reserve((size_t)n_terms);
  m.name_to_id.reserve((size_t)std::min<int64_t>(offsets[n_terms], 1 << 24));
  std::vector<int32_t> members;
  for (int64_t k = 0; k < n_terms; ++k) {
    members.clear();
    for (int64_t j = offsets[k]; j < offsets[k + 1]; ++j) {
      const int32_t name = ids[j];
      auto ins = m.name_to_id.emplace(name, (int32_t)m.names.size());
      if (ins.second) m.names.push_back(name);
      const int32_t id = ins.first->second;
      if (std::find(members.begin(), members.end(), id) == members.end()) {
        members.push_back(id);
      } else if (kind == QMC_ISING) {
        throw Error(kDuplicatedVariable, "Duplicated ids detected in term!");
      }
    }
    std::sort(members.begin(), members.end());
    const double c = coeffs[k];
    const uint32_t loc = (uint32_t)members.size();
    m.max_locality = std::max(m.max_locality, loc);
    m.min_locality = std::min(m.min_locality, loc);
    m.total_locality += loc;
    if (loc > 0) {
      m.dependent_vars += (uint64_t)loc * (loc - 1);
      const double a = std::fabs(c);
      m.max_coupling = std::max(m.max_coupling, a);
      m.min_coupling = std::min(m.min_coupling, a);
      m.term_var.insert(m.term_var.end(), members.begin(), members.end());
      m.term_off.push_back((int64_t)m.term_var.size());
      m.coeff.push_back(c);
    } else {
      m.const_cost += c;
    }
  }
  m.n_vars = (int32_t)m.names.size();
  return m;
}

Model make_model(int kind, const std::vector<Term>& terms) {
  std::vector<int64_t> off(1, 0);
  std::vector<int32_t> ids;
  std::vector<double> c;
  for (const Term& t : terms) {
    ids.insert(ids.end(), t.ids.begin(), t.ids.end());
    off.push_back((int64_t)ids.size());
    c.push_back(t.c);
  }
  return make_model_flat(kind, (int64_t)terms.size(), off.data(), ids.data(), c.data());
}

void apply_initial_value(Model& m, const std::string& name_text, int64_t v) {
  // GraphModelConfiguration::map_initial_configuration (model/graph_model.h:64-95)
  const int32_t name = (int32_t)std::stol(name_text);
  auto it = m.name_to_id.find(name);
  if (it == m.name_to_id.end())
    throw Error(kInitialConfig, "parameter `initial_configuration` have node name: " + name_text +
                                    " which is not found in model node names");
  uint8_t bit;
  if (m.kind == QMC_ISING) {
    if (v != 1 && v != -1)
      throw Error(kValueError, "parameter `initial_configuration` may have 1 and -1 values for ising models, found: " + std::to_string(v));
    bit = v == -1;
  } else {
    if (v != 0 && v != 1)
      throw Error(kValueError, "parameter `initial_configuration` may have 1 and 0 values for pubo models, found: " + std::to_string(v));
    bit = v == 0;
  }
  m.initial.emplace_back(it->second, bit);
}

void finish_initial_configuration(Model& m) {
  if (!m.initial.empty() && (int32_t)m.initial.size() != m.n_vars)
    throw Error(kInitialConfig, "parameter `initial_configuration` have size: " +
                                    std::to_string(m.initial.size()) +
                                    " which is different from model nodes size: " +
                                    std::to_string(m.n_vars));
  m.has_initial = !m.initial.empty();
}

Model parse_model(const Json& doc) {
  const Json* cf = &doc;
  if (doc.is_object() && doc.has("cost_function")) cf = &doc.at("cost_function");
  if (!cf->is_object() || !cf->has("type"))
    throw Error(kMissingInput, "The configuration must contain a `cost_function` entry.");
  const std::string& type = cf->at("type").as_string();
  int kind;
  if (type == "ising") kind = QMC_ISING;
  else if (type == "pubo") kind = QMC_PUBO;
  else throw Error(kValueError, "unsupported cost_function type: " + type);
  std::string version = cf->has("version") ? cf->at("version").as_string() : std::string("1.0");
  // AbstractIsing/AbstractPubo::match_version (model/ising.h:80-89, model/pubo.h:69-78)
  if (version != "1.0" && version != "1.1")
    throw Error(kValueError, "Expecting `version` equals 1.0 or 1.1, but found: " + version);
  if (!cf->has("terms")) throw Error(kMissingInput, "parameter `terms` is required");
  const Json& jt = cf->at("terms");
  std::vector<int64_t> off(1, 0);
  std::vector<int32_t> ids;
  std::vector<double> c;
  off.reserve(jt.size() + 1);
  c.reserve(jt.size());
  for (const Json& t : jt.items()) {
    c.push_back(t.at("c").as_double());
    for (const Json& id : t.at("ids").items()) ids.push_back((int32_t)id.as_int());
    off.push_back((int64_t)ids.size());
  }
  Model m = make_model_flat(kind, (int64_t)c.size(), off.data(), ids.data(), c.data());
  m.version = version;
  if (cf->has("initial_configuration")) {
    for (const auto& kv : cf->at("initial_configuration").members()) apply_initial_value(m, kv.first, kv.second.as_int());
    finish_initial_configuration(m);
  }
  return m;
}

Json Model::render(const uint8_t* bits) const {
  // IsingState::render / BinaryWithCounter::render with the id map
  // (model/ising.cc:16-24, model/pubo.h:240-249)
  Json s = Json::object();
  for (int32_t i = 0; i < n_vars; ++i) {
    const int v = kind == QMC_ISING ? (bits[i] ? -1 : 1) : (bits[i] ? 0 : 1);
    s.append(std::to_string(names[i]), Json(v));  // names are unique
  }
  return s;
}

double Model::cost(const uint8_t* bits) const {
  double total = 0.0;
  for (size_t j = 0; j < coeff.size(); ++j) {
    uint32_t x = 0;
    for (int64_t k = term_off[j]; k < term_off[j + 1]; ++k) {
      const uint32_t b = bits[term_var[k]] & 1u;
      x = kind == QMC_ISING ? (x ^ b) : (x | b);
    }
    if (kind == QMC_ISING) total += x ? -coeff[j] : coeff[j];
    else if (!x) total += coeff[j];
  }
  return total;
}

qmc_problem DeviceModel::handle(int visit) {
  const int k = visit == QMC_VISIT_INDEX_LEVELS ? 1 : 0;
  if (!p_[k]) {
    const Model& m = *m_;
    device_check(qmc_problem_create_visit(ctx_, m.kind, m.n_vars, m.n_terms(), m.term_off.data(),
                                          m.term_var.data(), m.coeff.data(), lanes_, visit, &p_[k]),
                 "qmc_problem_create_visit");
    device_check(qmc_problem_get_info(p_[k], &info_[k]), "qmc_problem_get_info");
  }
  return p_[k];
}

const qmc_problem_info& DeviceModel::info(int visit) {
  handle(visit);
  return info_[visit == QMC_VISIT_INDEX_LEVELS ? 1 : 0];
}

DeviceModel::~DeviceModel() {
  for (qmc_problem p : p_)
    if (p) qmc_problem_destroy(p);
}

namespace {

// Largest differencing (Karmarkar-Karp) over one multiset, as
// utils::least_diff_method (utils/utils.cc:43-53): replace the two largest
// values by their difference until one value is left.
double largest_differencing(std::multiset<double> s) {
  if (s.empty()) return 0.0;
  while (s.size() > 1) {
    auto a = std::prev(s.end());
    const double x = *a;
    s.erase(a);
    auto b = std::prev(s.end());
    const double y = *b;
    s.erase(b);
    s.insert(x - y);
  }
  return *s.begin();
}

// The two-set form (utils/utils.cc:55-100): the largest of each set are
// paired off, the remainder returns to the set it came from (a tie returns
// one of them to the smaller set), and the result is the least nonzero
// remainder seen; with one set empty, the other's least value.
double largest_differencing(std::multiset<double> a, std::multiset<double> b) {
  double least = std::numeric_limits<double>::max();
  if (a.empty() && !b.empty()) return *b.begin();
  if (!a.empty() && b.empty()) return *a.begin();
  while (!a.empty() && !b.empty()) {
    auto ia = std::prev(a.end());
    const double x = *ia;
    a.erase(ia);
    auto ib = std::prev(b.end());
    const double y = *ib;
    b.erase(ib);
    double d = x - y;
    if (d > 0) {
      a.insert(d);
    } else if (d < 0) {
      d = -d;
      b.insert(d);
    } else if (a.size() < b.size()) {
      a.insert(x);
    } else {
      b.insert(y);
    }
    if (least > d && d != 0) least = d;
  }
  return least;
}

}  // namespace

double Model::estimate_max_cost_diff() const {
  std::vector<double> sum(n_vars, 0.0);
  for (int64_t t = 0; t < n_terms(); ++t)
    for (int64_t k = term_off[t]; k < term_off[t + 1]; ++k) sum[term_var[k]] += std::fabs(coeff[t]);
  double mx = 0.0;
  for (double v : sum) mx = std::max(mx, v);
  return kind == QMC_ISING ? 2.0 * mx : mx;
}

double Model::estimate_min_cost_diff() const {
  const double none = std::numeric_limits<double>::max();
  std::vector<std::multiset<double>> pos(n_vars), neg(kind == QMC_ISING ? 0 : n_vars);
  for (int64_t t = 0; t < n_terms(); ++t)
    for (int64_t k = term_off[t]; k < term_off[t + 1]; ++k) {
      const int32_t v = term_var[k];
      if (kind == QMC_ISING) pos[v].insert(std::fabs(coeff[t]));
      else if (coeff[t] > 0) pos[v].insert(coeff[t]);
      else if (coeff[t] < 0) neg[v].insert(-coeff[t]);
    }
  double least = none;
  for (int32_t v = 0; v < n_vars; ++v) {
    const double d = kind == QMC_ISING ? largest_differencing(pos[v]) : largest_differencing(pos[v], neg[v]);
    if (d == 0) continue;
    least = std::min(least, d);
  }
  if (least == none) least = 0.1;  // MIN_DELTA_DEFAULT (model/graph_model.h:16)
  return kind == QMC_ISING ? 2.0 * least : least;
}

}  // namespace qiotk
