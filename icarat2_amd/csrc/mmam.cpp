// mmam.cpp -- host side of MMA with m constraints (mmam.hpp).
#include "mmam.hpp"

#include <cmath>
#include <string>

namespace icarat_hip {

DeviceMMAm::DeviceMMAm(long long n, int m, double a, double c, double d) : a_(a), c_(c), d_(d) {
  if (n < 1 || m < 1 || m > kMmaMaxM)
    throw Error(kInvalidArgument, "MMA: n >= 1 and 1 <= m <= " + std::to_string(kMmaMaxM));
  P_.n = n;
  P_.m = m;
  P_.iter = 0;
  HIPCHK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
  for (DevBuf<double> *b : {&x_, &df_, &low_, &upp_, &alpha_, &beta_, &p0_, &q0_, &xold1_, &xold2_,
                            &xmin_, &xmax_})
    b->alloc(n);
  for (DevBuf<double> *b : {&dg_, &pij_, &qij_}) b->alloc(n * m);
  part_.alloc((size_t)kMmamPassSums * mmam_blocks(n));
  st_.alloc(1);
  hst_.alloc(1);
  // the history starts at zero (the reference's vectors are value-initialised)
  HIPCHK(hipMemsetAsync(xold1_.p, 0, xold1_.bytes(), s_));
  HIPCHK(hipMemsetAsync(xold2_.p, 0, xold2_.bytes(), s_));
  HIPCHK(hipMemsetAsync(low_.p, 0, low_.bytes(), s_));
  HIPCHK(hipMemsetAsync(upp_.p, 0, upp_.bytes(), s_));
  HIPCHK(hipStreamSynchronize(s_));
  A_ = MmamArrays{low_, upp_, alpha_, beta_, p0_, q0_, pij_, qij_, xold1_, xold2_, xmin_, xmax_};
}

DeviceMMAm::~DeviceMMAm() {
  if (s_) {
    (void)hipStreamSynchronize(s_);
    (void)hipStreamDestroy(s_);
  }
}

void DeviceMMAm::set_params(double move, double albefa, double asyminit, double asymdec,
                            double asyminc) {
  P_.move = move;
  P_.albefa = albefa;
  P_.asyminit = asyminit;
  P_.asymdec = asymdec;
  P_.asyminc = asyminc;
}

int DeviceMMAm::update(double *x, const double *dfdx, const double *gx, const double *dgdx,
                       const double *xmin, const double *xmax) {
  const long long n = P_.n;
  const int m = P_.m;
  HIPCHK(hipMemcpyAsync(x_.p, x, n * sizeof(double), hipMemcpyHostToDevice, s_));
  HIPCHK(hipMemcpyAsync(df_.p, dfdx, n * sizeof(double), hipMemcpyHostToDevice, s_));
  HIPCHK(hipMemcpyAsync(dg_.p, dgdx, n * m * sizeof(double), hipMemcpyHostToDevice, s_));
  HIPCHK(hipMemcpyAsync(xmin_.p, xmin, n * sizeof(double), hipMemcpyHostToDevice, s_));
  HIPCHK(hipMemcpyAsync(xmax_.p, xmax, n * sizeof(double), hipMemcpyHostToDevice, s_));
  // GenSub (the iteration counter advances first, MMASolver.cpp:404-405) and
  // the history shift, then SolveDIP from lambda = c/2, mu = 1
  P_.iter += 1;
  launch_mmam_gensub(P_, x_, df_, dg_, A_, part_, s_);
  const double epsimin = std::sqrt((double)(n + m)) * 1e-9;
  launch_mmam_init(P_, gx, a_, c_, epsimin, part_, st_, s_);
  // at most 100 outer loops of at most 100 Newton steps, each one launch
  const int max_steps = 100 * 101 + 1;
  int launched = 0, chunk = 24;
  for (;;) {
    for (int k = 0; k < chunk; ++k) launch_mmam_step(P_, A_, x_, part_, st_, s_);
    launched += chunk;
    HIPCHK(hipMemcpyAsync(hst_.p, st_.p, sizeof(MmamState), hipMemcpyDeviceToHost, s_));
    HIPCHK(hipStreamSynchronize(s_));
    if (hst_.p->done) break;
    if (launched > max_steps) throw Error(kInternal, "MMA dual iteration did not end");
    chunk = 16;
  }
  HIPCHK(hipMemcpyAsync(x, x_.p, n * sizeof(double), hipMemcpyDeviceToHost, s_));
  HIPCHK(hipStreamSynchronize(s_));
  steps_ = hst_.p->steps;
  (void)d_;  // d enters the reference only as "fixed at one" (MMASolver.cpp:366-371)
  return hst_.p->outer;
}

void DeviceMMAm::asymptotes(double *low, double *upp) {
  if (low) HIPCHK(hipMemcpyAsync(low, low_.p, low_.bytes(), hipMemcpyDeviceToHost, s_));
  if (upp) HIPCHK(hipMemcpyAsync(upp, upp_.p, upp_.bytes(), hipMemcpyDeviceToHost, s_));
  HIPCHK(hipStreamSynchronize(s_));
}

}  // namespace icarat_hip
