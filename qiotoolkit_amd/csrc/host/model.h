// Ising / PUBO cost functions in the reference's JSON format, normalised as
// graph::normalize_edges does (graph/graph.h:74-177) and uploaded to the
// device through the C ABI.
#pragma once
#include <cstdint>
#include <map>
#include <optional>
#include <unordered_map>
#include <string>
#include <vector>

#include "../../../include/qmc.h"
#include "errors.h"
#include "json.h"

namespace qiotk {

struct Term {
  double c;
  std::vector<int32_t> ids;
};

// A cost function after normalisation: internal variable ids 0..N-1 by first
// appearance (this is the reference's sweep order), ids sorted inside each
// term, constant terms folded into const_cost; duplicate ids are an error for
// Ising and merged for PUBO (model/pubo.h:82).
struct Model {
  int kind = QMC_ISING;          // QMC_ISING or QMC_PUBO
  std::string version = "1.0";
  int32_t n_vars = 0;
  std::vector<int64_t> term_off;  // n_terms + 1
  std::vector<int32_t> term_var;
  std::vector<double> coeff;
  double const_cost = 0.0;
  std::vector<int32_t> names;     // internal id -> input id
  std::unordered_map<int32_t, int32_t> name_to_id;
  // graph properties (graph/graph_properties.h)
  uint32_t max_locality = 0, min_locality = UINT32_MAX;
  uint64_t total_locality = 0;
  uint64_t dependent_vars = 0;  // sum of locality x (locality - 1) over non-constant terms
  int64_t source_bytes = 0;     // size of the input the model was read from (0: term arrays)
  bool streamed = false;        // read by the streaming reader (else a tree parse or arrays)
  double max_coupling = 0.0, min_coupling = 1.7976931348623157e308;
  // initial_configuration, internal ids -> reference bit (true = -1 / 0)
  std::vector<std::pair<int32_t, uint8_t>> initial;
  bool has_initial = false;

  int64_t n_terms() const { return (int64_t)coeff.size(); }
  bool empty() const { return n_vars == 0; }
  // scale_factor used by population annealing (graph_properties.h:24-40)
  double scale_factor() const { return max_coupling <= 1 ? 1.0 : 1.0 / max_coupling; }
  // render a state (one byte per internal variable, 1 = reference true bit)
  Json render(const uint8_t* bits) const;
  // FP64 cost of a state (host-side, used for the const-cost-free checks)
  double cost(const uint8_t* bits) const;
  // Bounds on the cost change of one flip that set the parameter-free beta
  // range: the largest sum of |c| over a variable's terms (graph/graph.h:453-487),
  // and the least difference among a variable's coefficients by largest
  // differencing (utils/utils.cc:43-100), minimised over variables, 0.1 when
  // every variable's is 0; Ising doubles both (model/ising.h:176-224), PUBO
  // splits positive and negative coefficients (model/pubo.h:127-185).
  double estimate_max_cost_diff() const;
  double estimate_min_cost_diff() const;
};

// Parse {"cost_function": {...}} (or the inner object) and normalise it.
Model parse_model(const Json& doc);
Model make_model(int kind, const std::vector<Term>& terms);
// The same from a term CSR (offsets[n_terms + 1] into ids, one coefficient
// per term).
Model make_model_flat(int kind, int64_t n_terms, const int64_t* offsets, const int32_t* ids,
                      const double* coeffs);
// initial_configuration of a built model: each {input id: value} member in
// document order, then the size check
void apply_initial_value(Model& m, const std::string& name_text, int64_t value);
void finish_initial_configuration(Model& m);

// Streaming input (the reference reads its input through a SAX reader in
// 64 KiB file chunks, utils/stream_handler_json.h:124-182): the document is
// read once, straight into the term CSR, with no document tree.  nullopt
// when the document is not of the usual shape ({"cost_function": {"type",
// "version", "terms": [{"c", "ids"}], "initial_configuration"}} plus other
// keys, which are skipped) -- the caller then parses it as a tree, which
// gives the reference's error for a malformed document.
std::optional<Model> stream_model_text(const char* text, size_t n);
std::optional<Model> stream_model_file(const std::string& path);
// Parse a document from text or a file (a protobuf folder goes through the
// wire-format reader), streaming when the shape allows.
Model read_model_text(const char* text);
Model read_model_file(const std::string& path);

// Device copy of a model, owned.
// The model on the device, once per visit order (QMC_VISIT_COLOR_CLASSES,
// QMC_VISIT_INDEX_LEVELS), each uploaded when first asked for.  `m` must
// outlive it.
class DeviceModel {
 public:
  DeviceModel(qmc_ctx ctx, const Model& m, int lanes = 0) : ctx_(ctx), m_(&m), lanes_(lanes) {}
  ~DeviceModel();
  DeviceModel(const DeviceModel&) = delete;
  DeviceModel& operator=(const DeviceModel&) = delete;
  qmc_problem handle(int visit = QMC_VISIT_COLOR_CLASSES);
  const qmc_problem_info& info(int visit = QMC_VISIT_COLOR_CLASSES);

 private:
  qmc_ctx ctx_;
  const Model* m_;
  int lanes_;
  qmc_problem p_[2] = {nullptr, nullptr};
  qmc_problem_info info_[2]{};
};

}  // namespace qiotk
