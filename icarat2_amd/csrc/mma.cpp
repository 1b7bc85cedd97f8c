// mma.cpp -- host side of the device MMA (see mma.hpp).
#include "mma.hpp"

#include <algorithm>
#include <cmath>

#include "context.hpp"

namespace icarat_hip {

int solve_dual_m1(Context &ctx, long long n, const MmaArrays &A, double *x, double b, double a,
                  double c, int *newton_steps) {
  hipStream_t s = ctx.stream();
  double *sc = ctx.d_scal.p + 4;
  double *hs = ctx.h_scal.p + 4;
  // SolveDIP (reference MMASolver.cpp:88-154, GCMMASolver.cpp:128-196), m = 1
  const double epsimin = std::sqrt((double)(n + 1)) * 1e-9;
  double lam = c / 2.0, mu = 1.0;
  auto pass = [&](double l, double &Sg, double &Sh) {
    launch_mma_dual(n, l, A, x, ctx.d_partials, ctx.d_counter, sc, s);
    HIPCHK(hipMemcpyAsync(hs, sc, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    ctx.sync();
    Sg = hs[0];
    Sh = hs[1];
  };
  double Sg = 0.0, Sh = 0.0;
  pass(lam, Sg, Sh);
  double epsi = 1.0, err = 1.0;
  int outer = 0, steps = 0;
  while (epsi > epsimin && outer < 100) {
    ++outer;
    int loop = 0;
    while (err > 0.9 * epsi && loop < 100) {
      ++loop;
      ++steps;
      lam = std::max(lam, 0.0);
      double y = std::max(0.0, lam - c);
      double z = std::max(0.0, 10.0 * (lam * a - 1.0));
      double grad = -b - a * z - y + Sg;
      grad = -1.0 * grad - epsi / lam;
      double hess = Sh;
      if (lam > c) hess += -1.0;
      hess += -mu / lam;
      if (lam * a > 0.0) hess += -10.0 * a * a;
      double corr = 1e-4 * hess / 1.0;
      if (-1.0 * corr < 1.0e-7) corr = -1.0e-7;
      hess += corr;
      const double s0 = grad / hess;
      const double s1 = -mu + epsi / lam - s0 * mu / lam;
      // line search keeping lambda, mu positive (reference MMASolver.cpp:209-236)
      double th = 1.005;
      th = std::max(th, -1.01 * s0 / lam);
      th = std::max(th, -1.01 * s1 / mu);
      th = 1.0 / th;
      lam += th * s0;
      mu += th * s1;
      pass(lam, Sg, Sh);
      y = std::max(0.0, lam - c);
      z = std::max(0.0, 10.0 * (lam * a - 1.0));
      const double r1 = -b - a * z - y + mu + Sg;
      const double r2 = mu * lam - epsi;
      err = std::max(std::abs(r1), std::abs(r2));
    }
    epsi *= 0.1;
  }
  if (newton_steps) *newton_steps = steps;
  return outer;
}

DeviceMMA::DeviceMMA(Context &ctx) : ctx_(ctx), n_(ctx.lattice().nelx) {
  for (auto &b : buf_) {
    b.alloc(n_);
    HIPCHK(hipMemsetAsync(b.p, 0, b.bytes(), ctx_.stream()));
  }
  A_ = MmaArrays{buf_[0], buf_[1], buf_[2], buf_[3], buf_[4],
                 buf_[5], buf_[6], buf_[7], buf_[8], buf_[9]};
  P_.iter = 0;
  ctx_.sync();
}

void DeviceMMA::reset() {
  P_.iter = 0;
  HIPCHK(hipMemsetAsync(A_.xold1, 0, n_ * sizeof(double), ctx_.stream()));
  HIPCHK(hipMemsetAsync(A_.xold2, 0, n_ * sizeof(double), ctx_.stream()));
}

void DeviceMMA::set_params(double move, double albefa, double asyminit, double asymdec,
                           double asyminc) {
  P_.move = move;
  P_.albefa = albefa;
  P_.asyminit = asyminit;
  P_.asymdec = asymdec;
  P_.asyminc = asyminc;
}

int DeviceMMA::update(double *x, const double *dfdx, double g, const double *dgdx, double xmin,
                      double xmax) {
  hipStream_t s = ctx_.stream();
  double *sc = ctx_.d_scal.p + 4;
  double *hs = ctx_.h_scal.p + 4;
  // GenSub (the iteration counter advances first, reference MMASolver.cpp:404-405)
  P_.iter += 1;
  P_.xmin = xmin;
  P_.xmax = xmax;
  launch_mma_gensub(n_, x, dfdx, dgdx, P_, A_, ctx_.d_partials, ctx_.d_counter, sc, s);
  HIPCHK(hipMemcpyAsync(hs, sc, sizeof(double), hipMemcpyDeviceToHost, s));
  ctx_.sync();
  const double b = -g + hs[0];
  return solve_dual_m1(ctx_, n_, A_, x, b, a_, c_, &newton_steps_);
}

}  // namespace icarat_hip
