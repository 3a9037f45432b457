// PopulationAnnealing (solver/population_annealing.h) -- configuration; the
// stepping runs in solver_pa.cc.
#include "solver.h"

namespace qiotk {

void PopulationAnnealing::configure(const Json& doc) {
  Solver::configure(doc);
  throw Error(kNotImplemented, "populationannealing: not available yet");
}
void PopulationAnnealing::init() {}
void PopulationAnnealing::run() {}
void PopulationAnnealing::finalize() {}
PopulationAnnealing::~PopulationAnnealing() = default;

}  // namespace qiotk
