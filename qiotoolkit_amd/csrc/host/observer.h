// The reference's measurement protocol (observe/observer.h:14-129,
// observer.cc): solvers record named observations under labels ("step",
// "replica"); each configured protocol groups them (averaged over windows of
// a label relative to the last restart, then by labels) and keeps Constant
// (last value) or Average (weighted mean) observables; the tree renders into
// benchmark.extensions "solver".
#pragma once
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "json.h"

namespace qiotk {

class Observer {
 public:
  // {"observer": {...}} or {"observer": [{...}, ...]} at the top level of the
  // params document (Observer::configure, observer.cc:50-64)
  void configure(const Json& doc);
  bool watching(const std::string& name) const;
  bool empty() const { return protocols_.empty(); }
  void restart(int64_t step) { restart_step_ = step; }
  void set_label(const std::string& name, int64_t value) { labels_[name] = value; }
  void clear_label(const std::string& name) { labels_.erase(name); }
  void observe(const std::string& name, double value, double weight = 1.0);
  // `count` observations of weight 1 summing to `sum`, ending with `last`
  // (what a loop of observe(name, v_i) leaves, without the loop)
  void observe_many(const std::string& name, double sum, int64_t count, double last);
  Json render() const;

 private:
  enum Kind { kConstant, kAverage };
  struct Obs {
    Kind kind;
    double value = 0, sum = 0, weight = 0;
    bool set = false;
  };
  struct Group {
    std::map<std::string, std::map<int64_t, Group>> groups;
    std::map<std::string, Obs> obs;
    Json render() const;
  };
  struct Protocol {
    std::vector<std::pair<std::string, int64_t>> average_over;
    std::vector<std::string> group_by;
    std::map<std::string, Kind> kinds;
  };
  Group* select(const Protocol& p);
  std::vector<Protocol> protocols_;
  std::map<std::string, int64_t> labels_;
  Group root_;
  int64_t restart_step_ = 0;
};

}  // namespace qiotk
