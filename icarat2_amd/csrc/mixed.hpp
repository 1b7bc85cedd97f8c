// mixed.hpp -- FP32 / BF16 MFMA element contraction inside FP64 iterative
// refinement of K u = f.
#pragma once
#include "../../include/icarat_hip.h"
#include "hip_util.hpp"

namespace icarat_hip {

class Context;

class MixedSolver {
 public:
  explicit MixedSolver(Context &ctx);
  ~MixedSolver();
  void invalidate();  // densities changed
  // y = K u on the lattice with the FP32 or BF16 contraction (device pointers, FP64 vectors)
  void apply(const double *u, double *y, int precision);
  // Solve K u = ctx.d_b into ctx.d_x by FP64 refinement around a low-precision inner solve.
  void solve(const icarat_solve_opts &o, icarat_solve_stats &st);

 private:
  Context &ctx_;
};

}  // namespace icarat_hip
