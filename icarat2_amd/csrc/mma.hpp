// mma.hpp -- MMA design update with one (volume) constraint on the device:
// element passes and the 1x1 dual Newton system on the GPU, all FP64.
#pragma once
#include "hip_util.hpp"
#include "mma_kernels.hpp"

namespace icarat_hip {

class Context;

// The dual interior-point iteration of MMA / GCMMA for one constraint
// (SolveDIP, reference MMASolver.cpp:88-154, GCMMASolver.cpp:128-196): the
// element passes (x(lambda) and the dual sums) and the 1x1 Newton system, both
// on the device.  The constraint constant of the subproblem is b = b0, plus
// bsum[0] (device) when bsum is given; x receives x(lambda*).  Returns the
// outer (epsilon) loops.
int solve_dual_m1(Context &ctx, long long n, const MmaArrays &A, double *x, double b0,
                  const double *bsum, double a, double c, int *newton_steps);

class DeviceMMA {
 public:
  explicit DeviceMMA(Context &ctx);
  void reset();  // iteration counter and history back to the start
  void set_params(double move, double albefa, double asyminit, double asymdec, double asyminc);
  // MMASolver::Update for m = 1 (reference optimization/opt-solver/MMASolver.cpp:66-82),
  // a = 0, c = 1e5, d = 0.  x, dfdx, dgdx are device pointers; x updated in place.
  // Returns the number of outer interior-point loops.
  int update(double *x, const double *dfdx, double g, const double *dgdx, double xmin,
             double xmax);
  int last_newton_steps() const { return newton_steps_; }

 private:
  Context &ctx_;
  long long n_;
  DevBuf<double> buf_[10];
  MmaArrays A_;
  MmaParams P_;
  double a_ = 0.0, c_ = 1.0e5, d_ = 0.0;
  int newton_steps_ = 0;
};

}  // namespace icarat_hip
