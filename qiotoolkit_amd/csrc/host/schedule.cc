#include "schedule.h"

#include <cmath>

#include "errors.h"

namespace qiotk {

// ScheduleGenerator (schedule/schedule_generator.cc)
class Generator {
 public:
  std::optional<double> start_, stop_;
  std::optional<int> count_;
  bool repeat_ = false;
  virtual ~Generator() = default;
  virtual double progress_value(double p) const = 0;
  virtual std::unique_ptr<Generator> clone() const = 0;
  virtual Json render() const = 0;
  // a string that determines every value of the schedule
  virtual std::string key() const { return render().dump(); }
  virtual bool is_constant() const { return false; }
  double get_start() const { return start_ ? *start_ : 0.0; }
  double get_stop() const { return stop_ ? *stop_ : 1.0; }
  int get_count() const {
    return count_ ? *count_ : (int)std::round(std::fabs(get_stop() - get_start()));
  }
  void configure_base(const Json& j) {
    if (j.has("start")) start_ = j.at("start").as_double();
    if (j.has("count")) count_ = (int)j.at("count").as_int();
    if (j.has("stop")) stop_ = j.at("stop").as_double();
    if (j.has("repeat")) repeat_ = j.at("repeat").as_bool();
  }
  void render_base(Json& o) const {
    if (start_) o["start"] = *start_;
    if (stop_) o["stop"] = *stop_;
    if (count_) o["count"] = *count_;
    if (repeat_) o["repeat"] = true;
  }
};

namespace {

class Linear : public Generator {  // schedule/linear.cc
 public:
  double a = 0, b = 0;
  double progress_value(double p) const override { return a + p * (b - a); }
  std::unique_ptr<Generator> clone() const override { return std::make_unique<Linear>(*this); }
  Json render() const override {
    Json o = Json::object();
    o["type"] = "linear";
    o["initial"] = a;
    o["final"] = b;
    render_base(o);
    return o;
  }
};

class Geometric : public Generator {  // schedule/geometric.cc
 public:
  double a = 1, b = 1;
  double progress_value(double p) const override { return a * std::pow(b / a, p); }
  std::unique_ptr<Generator> clone() const override { return std::make_unique<Geometric>(*this); }
  Json render() const override {
    Json o = Json::object();
    o["type"] = "geometric";
    o["initial"] = a;
    o["final"] = b;
    render_base(o);
    return o;
  }
};

class Constant : public Generator {  // schedule/constant.cc
 public:
  double v = 0;
  double progress_value(double) const override { return v; }
  std::unique_ptr<Generator> clone() const override { return std::make_unique<Constant>(*this); }
  bool is_constant() const override { return true; }
  Json render() const override {
    Json o = Json::object();
    o["type"] = "constant";
    o["value"] = v;
    render_base(o);
    return o;
  }
};

}  // namespace

class Segments : public Generator {  // schedule/segments.cc
 public:
  std::vector<Schedule> segs;
  // each segment's adjusted start/stop (segments.cc:63-129)
  std::vector<double> seg_start, seg_stop;
  std::vector<bool> seg_point;

  double progress_value(double progress) const override {
    const double position = get_start() + progress * (get_stop() - get_start());
    if (position < seg_start.front()) return segs.front().get_initial();
    for (size_t i = 0; i < segs.size(); ++i) {
      const double stop = seg_stop[i];
      if (position > stop) continue;
      const double start = seg_start[i];
      if (i > 0 && position < start) {
        const double last_stop = seg_stop[i - 1];
        const double last_value = segs[i - 1].get_final();
        const double next_value = segs[i].get_initial();
        const double pr = (position - last_stop) / (start - last_stop);
        return (1 - pr) * last_value + pr * next_value;
      }
      return segs[i].get_value(position);
    }
    return segs.back().get_final();
  }
  std::unique_ptr<Generator> clone() const override { return std::make_unique<Segments>(*this); }
  Json render() const override {
    Json o = Json::object();
    o["type"] = "segments";
    Json a = Json::array();
    for (const auto& s : segs) a.push_back(s.render());
    o["segments"] = a;
    return o;
  }
  // the rendering leaves out the outer start/stop/count/repeat
  std::string key() const override {
    Json o = Json::object();
    o["type"] = "segments";
    render_base(o);
    std::string k = o.dump();
    for (const auto& s : segs) k += "|" + s.key();
    return k;
  }
};

Schedule::Schedule() = default;
Schedule::~Schedule() = default;
Schedule::Schedule(const Schedule& o) : g_(o.g_ ? o.g_->clone() : nullptr) {}
Schedule& Schedule::operator=(const Schedule& o) {
  if (this != &o) g_ = o.g_ ? o.g_->clone() : nullptr;
  return *this;
}

namespace {

bool type_is(const Json& j, const char* t) {
  return j.is_object() && j.has("type") && j.at("type").is_string() && j.at("type").as_string() == t;
}

}  // namespace

void Schedule::configure(const Json& j) {
  if (type_is(j, "linear") || type_is(j, "geometric")) {
    if (!j.has("initial") || !j.has("final"))
      throw Error(kMissingInput, "schedule: `initial` and `final` are required");
    if (type_is(j, "linear")) {
      auto g = std::make_unique<Linear>();
      g->configure_base(j);
      g->a = j.at("initial").as_double();
      g->b = j.at("final").as_double();
      g_ = std::move(g);
    } else {
      auto g = std::make_unique<Geometric>();
      g->configure_base(j);
      g->a = j.at("initial").as_double();
      g->b = j.at("final").as_double();
      if (g->a <= 0) throw Error(kValueError, "initial must be positive for geometric");
      if (g->b <= 0) throw Error(kValueError, "final must be positive for geometric");
      g_ = std::move(g);
    }
    return;
  }
  if (type_is(j, "segments") || j.is_array()) {
    const Json& arr = j.is_array() ? j : j.at("segments");
    if (!arr.is_array() || arr.size() == 0)
      throw Error(kValueError, "`segments` must be specified as a non-empty array.");
    auto g = std::make_unique<Segments>();
    if (j.is_object()) g->configure_base(j);
    for (const Json& e : arr.items()) {
      Schedule s;
      s.configure(e);
      const bool point = s.g_->is_constant() && !s.g_->stop_ && !s.g_->count_;
      g->seg_point.push_back(point);
      g->segs.push_back(std::move(s));
    }
    // Segments::adjust (segments.cc:63-129)
    const size_t n = g->segs.size();
    g->seg_start.assign(n, 0);
    g->seg_stop.assign(n, 0);
    double last_stop = g->segs[0].g_->get_start();
    for (size_t i = 0; i < n; ++i) {
      Generator* cur = g->segs[i].g_.get();
      if (!cur->start_) cur->start_ = last_stop;
      if (i > 0 && cur->get_start() < g->segs[i - 1].g_->get_stop())
        throw Error(kValueError, "Overlapping segments");
      if (g->seg_point[i]) {
        cur->stop_ = cur->get_start();
        last_stop = cur->get_stop() + 1;
      } else if (!cur->stop_) {
        const int count = cur->get_count();
        cur->stop_ = cur->get_start() + count;
        last_stop = cur->get_stop();
      }
      if (cur->get_stop() < cur->get_start()) throw Error(kValueError, "Negative width segment");
      g->seg_start[i] = cur->get_start();
      g->seg_stop[i] = cur->get_stop();
    }
    if (!g->start_) g->start_ = g->seg_start.front();
    if (!g->stop_) g->stop_ = g->seg_stop.back();
    if (!g->count_) {
      const int fencepost = (int)std::floor(last_stop - g->seg_stop.back());
      g->count_ = (int)std::round(g->get_stop() - g->get_start() + fencepost);
    }
    g_ = std::move(g);
    return;
  }
  if (j.is_number() || type_is(j, "constant") || (j.is_object() && !j.has("type") && j.has("value"))) {
    auto g = std::make_unique<Constant>();
    if (j.is_number()) {
      g->v = j.as_double();
    } else {
      g->configure_base(j);
      if (!j.has("value")) throw Error(kMissingInput, "schedule: `value` is required");
      g->v = j.at("value").as_double();
    }
    g_ = std::move(g);
    return;
  }
  throw Error(kInitialConfig, "unable to initialize schedule");
}

void Schedule::make_linear(double a, double b) {
  auto g = std::make_unique<Linear>();
  g->a = a;
  g->b = b;
  g_ = std::move(g);
}

void Schedule::make_geometric(double a, double b) {
  auto g = std::make_unique<Geometric>();
  g->a = a;
  g->b = b;
  g_ = std::move(g);
}

void Schedule::make_list(const std::vector<double>& values) {
  Json arr = Json::array();
  for (double v : values) arr.push_back(Json(v));
  configure(arr);
}

void Schedule::make_constant(double v) {
  auto g = std::make_unique<Constant>();
  g->v = v;
  g_ = std::move(g);
}

double Schedule::get_value(double input) const {
  const double start = g_->get_start(), stop = g_->get_stop();
  if (start == stop) return g_->progress_value(0);
  double progress = (input - start) / (stop - start);
  if (g_->repeat_) progress -= std::floor(progress);
  return g_->progress_value(std::min(1., std::max(0., progress)));
}

std::vector<double> Schedule::get_discretized_values(int count) const {
  if (count == -1 && g_->count_) count = *g_->count_;
  if (count == -1) count = (int)std::round(std::fabs(g_->get_stop() - g_->get_start()) + 1);
  if (count < 2) throw Error(kValueError, "Can not get_discretized_values(count) for less than 2 values");
  std::vector<double> out;
  out.reserve(count);
  for (int i = 0; i < count; ++i) out.push_back(g_->progress_value(double(i) / (count - 1)));
  return out;
}

double Schedule::get_initial() const { return g_->progress_value(0.0); }
double Schedule::get_final() const { return g_->progress_value(1.0); }
double Schedule::get_start() const { return g_->get_start(); }
double Schedule::get_stop() const { return g_->get_stop(); }
int Schedule::get_count() const { return g_->get_count(); }
void Schedule::set_stop(int stop) {
  if (g_) g_->stop_ = (double)stop;
}
std::string Schedule::key() const { return g_ ? g_->key() : std::string("none"); }

Json Schedule::render() const { return g_ ? g_->render() : Json(); }

}  // namespace qiotk
