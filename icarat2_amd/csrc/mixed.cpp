// mixed.cpp -- mixed-precision solver (see mixed.hpp).
#include "mixed.hpp"

#include "context.hpp"

namespace icarat_hip {

MixedSolver::MixedSolver(Context &ctx) : ctx_(ctx) {}
MixedSolver::~MixedSolver() = default;
void MixedSolver::invalidate() {}

void MixedSolver::apply(const double *, double *, int) {
  throw Error(kInvalidArgument, "FP32/BF16 contraction is not available in this build");
}

void MixedSolver::solve(const icarat_solve_opts &, icarat_solve_stats &) {
  throw Error(kInvalidArgument, "mixed-precision and learned solves are not available in this build");
}

}  // namespace icarat_hip
