// gcmma.cpp -- host side of the device GCMMA (see gcmma.hpp).
#include "gcmma.hpp"

#include <algorithm>
#include <cmath>

#include "context.hpp"
#include "mma.hpp"

namespace icarat_hip {

namespace {
constexpr double kRaaEps = 1.0e-6;  // raa0eps = raaeps (GCMMASolver.cpp:41-44)
}

DeviceGCMMA::DeviceGCMMA(Context &ctx) : ctx_(ctx), n_(ctx.lattice().nelx) {
  for (auto &b : buf_) b.alloc(n_);
  A_ = MmaArrays{buf_[0], buf_[1], buf_[2], buf_[3], buf_[4],
                 buf_[5], buf_[6], buf_[7], buf_[8], buf_[9]};
  epsimin_ = std::sqrt((double)(n_ + 1)) * 1e-9;
  reset();
}

void DeviceGCMMA::reset() {
  P_.iter = 0;
  raa0_ = 1.0;
  raa_ = 1.0;
  f0app_ = fapp_ = 0.0;
  for (auto &b : buf_) HIPCHK(hipMemsetAsync(b.p, 0, b.bytes(), ctx_.stream()));
  ctx_.sync();
}

void DeviceGCMMA::set_params(double move, double albefa, double asyminit, double asymdec,
                             double asyminc) {
  P_.move = move;
  P_.albefa = albefa;
  P_.asyminit = asyminit;
  P_.asymdec = asymdec;
  P_.asyminc = asyminc;
}

// GenSub with the constants r0 = f0x - sum(...), r1 = fx - sum(...), b = -r1
void DeviceGCMMA::gensub(const double *xval, double f0x, const double *df0dx, double fx,
                         const double *dfdx, int asymp) {
  hipStream_t s = ctx_.stream();
  double *sc = ctx_.d_scal.p + 4;
  double *hs = ctx_.h_scal.p + 4;
  launch_gcmma_sub(n_, xval, df0dx, dfdx, P_, raa0_, raa_, asymp, A_, ctx_.d_partials, sc, s);
  HIPCHK(hipMemcpyAsync(hs, sc, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  ctx_.sync();
  r0_ = f0x - hs[0];
  r1_ = fx - hs[1];
  b_ = -r1_;
}

// SolveDIP from xval, then ComputeApprox at the result
void DeviceGCMMA::solve_and_approx(double *xmma, const double *xval) {
  hipStream_t s = ctx_.stream();
  HIPCHK(hipMemcpyAsync(xmma, xval, n_ * sizeof(double), hipMemcpyDeviceToDevice, s));
  solve_dual_m1(ctx_, n_, A_, xmma, b_, nullptr, a_, c_, nullptr);
  double *sc = ctx_.d_scal.p + 4;
  double *hs = ctx_.h_scal.p + 4;
  launch_gcmma_approx(n_, xmma, A_, ctx_.d_partials, sc, s);
  HIPCHK(hipMemcpyAsync(hs, sc, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  ctx_.sync();
  f0app_ = hs[0] + r0_;
  fapp_ = hs[1] + r1_;
}

void DeviceGCMMA::outer(double *xmma, const double *xval, double f0x, const double *df0dx,
                        double fx, const double *dfdx, double xmin, double xmax) {
  if (!(xmax > xmin)) throw Error(kInvalidArgument, "GCMMA needs xmax > xmin");
  P_.xmin = xmin;
  P_.xmax = xmax;
  // Asymp: the outer counter advances, raa0 / raa shrink (GCMMASolver.cpp:486-560)
  P_.iter += 1;
  raa0_ = std::max(kRaaEps, 0.1 * raa0_);
  raa_ = std::max(kRaaEps, 0.1 * raa_);
  gensub(xval, f0x, df0dx, fx, dfdx, 1);
  solve_and_approx(xmma, xval);
}

void DeviceGCMMA::inner(double *xmma, double f0xnew, double fxnew, const double *xval,
                        double f0x, const double *df0dx, double fx, const double *dfdx,
                        double xmin, double xmax) {
  if (P_.iter == 0) throw Error(kNotReady, "GCMMA inner update before an outer update");
  P_.xmin = xmin;
  P_.xmax = xmax;
  // RaaUpdate (GCMMASolver.cpp:562-591)
  hipStream_t s = ctx_.stream();
  double *sc = ctx_.d_scal.p + 4;
  double *hs = ctx_.h_scal.p + 4;
  const double winv = 1.0 / std::max(P_.xmamieps, xmax - xmin);
  launch_gcmma_raacof(n_, xmma, xval, winv, A_, ctx_.d_partials, sc, s);
  HIPCHK(hipMemcpyAsync(hs, sc, sizeof(double), hipMemcpyDeviceToHost, s));
  ctx_.sync();
  const double cof = std::max(1e-12, hs[0]);
  if (f0xnew > f0app_ + 0.5 * epsimin_)
    raa0_ = std::min(1.1 * (raa0_ + (1.0 / cof) * (f0xnew - f0app_)), 10.0 * raa0_);
  if (fxnew > fapp_ + 0.5 * epsimin_)
    raa_ = std::min(1.1 * (raa_ + (1.0 / cof) * (fxnew - fapp_)), 10.0 * raa_);
  gensub(xval, f0x, df0dx, fx, dfdx, 0);
  solve_and_approx(xmma, xval);
}

bool DeviceGCMMA::concheck(double f0xnew, double fxnew) const {
  if (f0app_ + epsimin_ < f0xnew) return false;
  if (fapp_ + epsimin_ < fxnew) return false;
  return true;
}

void DeviceGCMMA::state(double out[4]) const {
  out[0] = raa0_;
  out[1] = raa_;
  out[2] = f0app_;
  out[3] = fapp_;
}

}  // namespace icarat_hip
