// mixed.hpp -- FP32 / BF16 MFMA element contraction inside FP64 iterative
// refinement of K u = f.
#pragma once
#include "../../include/icarat_hip.h"
#include "device_common.hpp"
#include "hip_util.hpp"

namespace icarat_hip {

class Context;

class MixedSolver {
 public:
  explicit MixedSolver(Context &ctx);
  ~MixedSolver();
  void invalidate() { ready_ = false; }  // densities changed
  // y = K u on the lattice with the FP32 or BF16 contraction (device pointers, FP64 vectors)
  void apply(const double *u, double *y, int precision);
  // Solve K u = ctx.d_b into ctx.d_x: FP64 refinement r = b - K64 x around an
  // FP32-storage Jacobi-PCG whose K.p runs on the matrix cores.
  void solve(const icarat_solve_opts &o, icarat_solve_stats &st);

  // inner-solve pieces reused by the learned preconditioner
  void prepare();
  float *r32() { return r32_.p; }
  float *dinv32() { return dinv32_.p; }

 private:
  Context &ctx_;
  long long n3_;
  bool ready_ = false;
  DevBuf<float> ke32_, young32_, dinv32_, r32_, p32_[2], q32_, d32_, u32_, y32_;
  // [mode][parity of the p buffers]
  hipGraph_t graph_[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  hipGraphExec_t exec_[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  static constexpr int kBatch = 16;
  void build_graph(int mode);
  void iterate(int mode, int maxit);
};

}  // namespace icarat_hip
