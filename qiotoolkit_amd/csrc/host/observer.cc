#include "observer.h"

#include <cmath>

#include "errors.h"

namespace qiotk {

namespace {

std::vector<std::string> names_of(const Json& a, const char* key) {
  if (!a.is_array()) throw Error(kInvalidTypes, std::string("observer `") + key + "` must be an array");
  std::vector<std::string> out;
  for (size_t i = 0; i < a.size(); ++i) out.push_back(a[i].as_string());
  return out;
}

// NaN (nothing recorded) has no JSON spelling: null
Json number(double v) { return std::isnan(v) ? Json() : Json(v); }

}  // namespace

// Observer::configure and Protocol::configure (observer.cc:50-64, 108-140):
// label maps and sets are iterated in name order, as the reference's
// std::map / std::set are; a name listed in both is an Average.
void Observer::configure(const Json& doc) {
  protocols_.clear();
  if (!doc.is_object() || !doc.has("observer")) return;
  const Json& o = doc.at("observer");
  std::vector<const Json*> list;
  if (o.is_array()) {
    for (size_t i = 0; i < o.size(); ++i) list.push_back(&o[i]);
  } else {
    list.push_back(&o);
  }
  for (const Json* pj : list) {
    const Json& j = *pj;
    if (!j.is_object()) throw Error(kInvalidTypes, "Expected object");
    Protocol p;
    if (j.has("average_over")) {
      const Json& a = j.at("average_over");
      if (!a.is_object() || a.members().empty())
        throw Error(kValueError, "parameter `average_over`: must not be empty");
      std::map<std::string, int64_t> m;
      for (const auto& [k, v] : a.members()) {
        const int64_t w = v.as_int();
        if (w <= 0) throw Error(kValueError, "parameter `average_over`: windows must be positive");
        m[k] = w;
      }
      p.average_over.assign(m.begin(), m.end());
    }
    if (j.has("group_by")) {
      const std::vector<std::string> g = names_of(j.at("group_by"), "group_by");
      if (g.empty()) throw Error(kValueError, "parameter `group_by`: must not be empty");
      const std::set<std::string> sorted(g.begin(), g.end());
      p.group_by.assign(sorted.begin(), sorted.end());
    }
    if (j.has("constant"))
      for (const std::string& n : names_of(j.at("constant"), "constant")) p.kinds[n] = kConstant;
    if (j.has("average"))
      for (const std::string& n : names_of(j.at("average"), "average")) p.kinds[n] = kAverage;
    protocols_.push_back(std::move(p));
  }
}

bool Observer::watching(const std::string& name) const {
  for (const Protocol& p : protocols_)
    if (p.kinds.count(name)) return true;
  return false;
}

// Protocol::select_group (observer.cc:148-178): windows of average_over
// labels start at the last restart; missing labels are skipped.
Observer::Group* Observer::select(const Protocol& p) {
  Group* g = &root_;
  for (const auto& [name, window] : p.average_over) {
    auto it = labels_.find(name);
    if (it == labels_.end()) continue;
    const int64_t v = it->second;
    g = &g->groups[name][v - (v - restart_step_) % window];
  }
  for (const std::string& name : p.group_by) {
    auto it = labels_.find(name);
    if (it == labels_.end()) continue;
    g = &g->groups[name][it->second];
  }
  return g;
}

void Observer::observe(const std::string& name, double value, double weight) {
  for (const Protocol& p : protocols_) {
    auto k = p.kinds.find(name);
    if (k == p.kinds.end()) continue;
    Group* g = select(p);
    auto ins = g->obs.try_emplace(name, Obs{k->second});
    Obs& o = ins.first->second;
    if (o.kind == kConstant) {  // Constant::record (observable.cc:21-27)
      if (weight > 0) {
        o.value = value;
        o.set = true;
      }
    } else {  // Average::record (observable.cc:35-39)
      o.weight += weight;
      o.sum += weight * value;
    }
  }
}

void Observer::observe_many(const std::string& name, double sum, int64_t count, double last) {
  if (count <= 0) return;
  for (const Protocol& p : protocols_) {
    auto k = p.kinds.find(name);
    if (k == p.kinds.end()) continue;
    Group* g = select(p);
    auto ins = g->obs.try_emplace(name, Obs{k->second});
    Obs& o = ins.first->second;
    if (o.kind == kConstant) {
      o.value = last;
      o.set = true;
    } else {
      o.weight += (double)count;
      o.sum += sum;
    }
  }
}

// Group::render (observer.cc:89-106): sub-groups as arrays in label-value
// order, then the observables.
Json Observer::Group::render() const {
  Json s = Json::object();
  for (const auto& [name, by_value] : groups) {
    Json a = Json::array();
    for (const auto& kv : by_value) a.push_back(kv.second.render());
    s[name] = a;
  }
  for (const auto& [name, o] : obs)
    s[name] = o.kind == kConstant ? number(o.set ? o.value : NAN) : number(o.weight > 0 ? o.sum / o.weight : NAN);
  return s;
}

Json Observer::render() const { return root_.render(); }

}  // namespace qiotk
