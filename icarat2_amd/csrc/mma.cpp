// mma.cpp -- host side of the device MMA (see mma.hpp).
#include "mma.hpp"

#include <algorithm>
#include <cmath>

#include "context.hpp"

namespace icarat_hip {

int solve_dual_m1(Context &ctx, long long n, const MmaArrays &A, double *x, double b0,
                  const double *bsum, double a, double c, int *newton_steps) {
  hipStream_t s = ctx.stream();
  MmaDualState *st = ctx.mma_dual_state();
  MmaDualState *hst = ctx.mma_dual_state_host();
  // SolveDIP (reference MMASolver.cpp:88-154, GCMMASolver.cpp:128-196), m = 1:
  // every pass and Newton step is enqueued on the device; the host only reads
  // the state after a chunk of passes to see whether the iteration has ended
  // (passes enqueued after the end return at once).
  const double epsimin = std::sqrt((double)(n + 1)) * 1e-9;
  launch_mma_dual_init(st, b0, bsum, a, c, epsimin, s);
  int chunk = 24;
  for (;;) {
    for (int k = 0; k < chunk; ++k) launch_mma_dual_step(n, A, x, ctx.d_partials, st, s);
    HIPCHK(hipMemcpyAsync(hst, st, sizeof(MmaDualState), hipMemcpyDeviceToHost, s));
    ctx.sync();
    if (hst->done) break;
    chunk = 16;
  }
  if (newton_steps) *newton_steps = hst->steps;
  return hst->outer;
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
  // GenSub (the iteration counter advances first, reference MMASolver.cpp:404-405)
  P_.iter += 1;
  P_.xmin = xmin;
  P_.xmax = xmax;
  launch_mma_gensub(n_, x, dfdx, dgdx, P_, A_, ctx_.d_partials, ctx_.d_counter, sc, s);
  // b = -g + sum p1/(U-x) + q1/(x-L), formed on the device
  return solve_dual_m1(ctx_, n_, A_, x, -g, sc, a_, c_, &newton_steps_);
}

}  // namespace icarat_hip
