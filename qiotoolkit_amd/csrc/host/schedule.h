// Beta / temperature schedules (cpp/schedule): Linear, Geometric, Constant
// and Segments, with the reference's start/stop/count/repeat semantics.
#pragma once
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "json.h"

namespace qiotk {

class Generator;

class Schedule {
 public:
  Schedule();
  Schedule(const Schedule& o);
  Schedule& operator=(const Schedule& o);
  ~Schedule();

  // Schedule::configure: try Linear, Geometric, Segments, Constant in turn
  // (schedule/schedule.cc:96-118).
  void configure(const Json& j);
  void make_linear(double a, double b);
  void make_geometric(double a, double b);
  void make_list(const std::vector<double>& values);
  void make_constant(double v);

  double get_value(double input) const;                         // schedule.cc:17-25
  std::vector<double> get_discretized_values(int count = -1) const;  // schedule.cc:38-89
  double get_initial() const;
  double get_final() const;
  double get_start() const;
  double get_stop() const;
  int get_count() const;
  void set_stop(int stop);
  Json render() const;
  // identifies the schedule's values (equal keys: equal values)
  std::string key() const;

 private:
  std::unique_ptr<Generator> g_;
};

}  // namespace qiotk
