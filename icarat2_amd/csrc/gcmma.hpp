// gcmma.hpp -- GCMMA design update with one (volume) constraint on the device
// (reference optimization/opt-solver/GCMMASolver.cpp): element passes on the
// GPU, the scalars (raa0, raa, the approximation constants, the 1x1 dual
// Newton system) on the host, all FP64.
#pragma once
#include "hip_util.hpp"
#include "mma_kernels.hpp"

namespace icarat_hip {

class Context;

class DeviceGCMMA {
 public:
  explicit DeviceGCMMA(Context &ctx);
  // back to the constructor's state: outer iteration 0, raa0 = raa = 1, no history
  void reset();
  void set_params(double move, double albefa, double asyminit, double asymdec, double asyminc);
  // GCMMASolver::OuterUpdate / InnerUpdate / ConCheck (GCMMASolver.hpp:69-85) for
  // m = 1 with a = 0, c = 1000, d = 0.  Vectors are device pointers of nelx;
  // xmma receives the new design.  Bounds are uniform.
  void outer(double *xmma, const double *xval, double f0x, const double *df0dx, double fx,
             const double *dfdx, double xmin, double xmax);
  void inner(double *xmma, double f0xnew, double fxnew, const double *xval, double f0x,
             const double *df0dx, double fx, const double *dfdx, double xmin, double xmax);
  bool concheck(double f0xnew, double fxnew) const;
  // {raa0, raa, f0app, fapp}
  void state(double out[4]) const;

 private:
  void gensub(const double *xval, double f0x, const double *df0dx, double fx, const double *dfdx,
              int asymp);
  void solve_and_approx(double *xmma, const double *xval);

  Context &ctx_;
  long long n_;
  DevBuf<double> buf_[10];
  MmaArrays A_;
  MmaParams P_;
  double a_ = 0.0, c_ = 1000.0;
  double raa0_ = 1.0, raa_ = 1.0, r0_ = 0.0, r1_ = 0.0, b_ = 0.0;
  double f0app_ = 0.0, fapp_ = 0.0, epsimin_ = 0.0;
};

}  // namespace icarat_hip
