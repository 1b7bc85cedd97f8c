// mmam.cpp -- host side of MMA with m constraints (mmam.hpp).
#include "mmam.hpp"

#include <algorithm>
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

// ---- GCMMA ------------------------------------------------------------------
namespace {
constexpr double kRaaEps = 1.0e-6;  // raa0eps = raaeps (GCMMASolver.cpp:41-44)
}

DeviceGCMMAm::DeviceGCMMAm(long long n, int m, double a, double c, double d)
    : a_(a), c_(c), d_(d) {
  if (n < 1 || m < 1 || m > kMmaMaxM)
    throw Error(kInvalidArgument, "GCMMA: n >= 1 and 1 <= m <= " + std::to_string(kMmaMaxM));
  P_.n = n;
  P_.m = m;
  P_.iter = 0;
  epsimin_ = std::sqrt((double)(n + m)) * 1e-9;
  for (int j = 0; j < kMmaMaxM; ++j) {
    raa_.r[j] = 1.0;
    r_.r[j] = 0.0;
    fapp_.r[j] = 0.0;
  }
  HIPCHK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
  for (DevBuf<double> *b : {&xv_, &xm_, &df_, &low_, &upp_, &alpha_, &beta_, &p0_, &q0_, &xold1_,
                            &xold2_, &xmin_, &xmax_})
    b->alloc(n);
  for (DevBuf<double> *b : {&dg_, &pij_, &qij_}) b->alloc(n * m);
  part_.alloc((size_t)kMmamPassSums * mmam_blocks(n));
  sums_.alloc(1 + kMmaMaxM);
  st_.alloc(1);
  hst_.alloc(1);
  hsum_.alloc(1 + kMmaMaxM);
  for (DevBuf<double> *b : {&xold1_, &xold2_, &low_, &upp_})
    HIPCHK(hipMemsetAsync(b->p, 0, b->bytes(), s_));
  HIPCHK(hipStreamSynchronize(s_));
  A_ = MmamArrays{low_, upp_, alpha_, beta_, p0_, q0_, pij_, qij_, xold1_, xold2_, xmin_, xmax_};
}

DeviceGCMMAm::~DeviceGCMMAm() {
  if (s_) {
    (void)hipStreamSynchronize(s_);
    (void)hipStreamDestroy(s_);
  }
}

void DeviceGCMMAm::set_params(double move, double albefa, double asyminit, double asymdec,
                              double asyminc) {
  P_.move = move;
  P_.albefa = albefa;
  P_.asyminit = asyminit;
  P_.asymdec = asymdec;
  P_.asyminc = asyminc;
}

void DeviceGCMMAm::load(const double *xval, const double *df0dx, const double *dfdx,
                        const double *xmin, const double *xmax) {
  const long long n = P_.n;
  HIPCHK(hipMemcpyAsync(xv_.p, xval, n * sizeof(double), hipMemcpyHostToDevice, s_));
  HIPCHK(hipMemcpyAsync(df_.p, df0dx, n * sizeof(double), hipMemcpyHostToDevice, s_));
  HIPCHK(hipMemcpyAsync(dg_.p, dfdx, n * P_.m * sizeof(double), hipMemcpyHostToDevice, s_));
  HIPCHK(hipMemcpyAsync(xmin_.p, xmin, n * sizeof(double), hipMemcpyHostToDevice, s_));
  HIPCHK(hipMemcpyAsync(xmax_.p, xmax, n * sizeof(double), hipMemcpyHostToDevice, s_));
}

// GenSub's constants: r0 = f0x - sum, r_j = fx_j - sum_j (b_j = -r_j)
void DeviceGCMMAm::subproblem(int asymp, double f0x, const double *fx) {
  const int m = P_.m;
  launch_gcmmam_sub(P_, asymp, xv_, df_, dg_, raa0_, raa_, A_, part_, sums_, s_);
  HIPCHK(hipMemcpyAsync(hsum_.p, sums_.p, (1 + m) * sizeof(double), hipMemcpyDeviceToHost, s_));
  HIPCHK(hipStreamSynchronize(s_));
  r0_ = f0x - hsum_.p[0];
  for (int j = 0; j < m; ++j) r_.r[j] = fx[j] - hsum_.p[1 + j];
}

// SolveDIP (GCMMASolver.cpp:135-197), then ComputeApprox (:598-620) at its x
void DeviceGCMMAm::solve_approx(double *xmma) {
  const long long n = P_.n;
  const int m = P_.m;
  MmamState S{};
  S.m = m;
  for (int j = 0; j < m; ++j) {
    S.b[j] = -r_.r[j];
    S.a[j] = a_;
    S.c[j] = c_;
    S.lam[j] = c_ / 2.0;
    S.mu[j] = 1.0;
  }
  S.epsi = 1.0;
  S.err = 1.0;
  S.epsimin = epsimin_;
  *hst_.p = S;
  HIPCHK(hipMemcpyAsync(st_.p, hst_.p, sizeof(MmamState), hipMemcpyHostToDevice, s_));
  HIPCHK(hipMemcpyAsync(xm_.p, xv_.p, n * sizeof(double), hipMemcpyDeviceToDevice, s_));
  const int max_steps = 100 * 101 + 1;
  int launched = 0, chunk = 24;
  for (;;) {
    for (int k = 0; k < chunk; ++k) launch_mmam_step(P_, A_, xm_, part_, st_, s_);
    launched += chunk;
    HIPCHK(hipMemcpyAsync(hst_.p, st_.p, sizeof(MmamState), hipMemcpyDeviceToHost, s_));
    HIPCHK(hipStreamSynchronize(s_));
    if (hst_.p->done) break;
    if (launched > max_steps) throw Error(kInternal, "GCMMA dual iteration did not end");
    chunk = 16;
  }
  launch_gcmmam_approx(P_, xm_, A_, part_, sums_, s_);
  HIPCHK(hipMemcpyAsync(hsum_.p, sums_.p, (1 + m) * sizeof(double), hipMemcpyDeviceToHost, s_));
  HIPCHK(hipMemcpyAsync(xmma, xm_.p, n * sizeof(double), hipMemcpyDeviceToHost, s_));
  HIPCHK(hipStreamSynchronize(s_));
  f0app_ = hsum_.p[0] + r0_;
  for (int j = 0; j < m; ++j) fapp_.r[j] = hsum_.p[1 + j] + r_.r[j];
}

void DeviceGCMMAm::outer(double *xmma, const double *xval, double f0x, const double *df0dx,
                         const double *fx, const double *dfdx, const double *xmin,
                         const double *xmax) {
  load(xval, df0dx, dfdx, xmin, xmax);
  // Asymp: the outer counter advances, raa0 / raa shrink (GCMMASolver.cpp:432-508)
  P_.iter += 1;
  raa0_ = std::max(kRaaEps, 0.1 * raa0_);
  for (int j = 0; j < P_.m; ++j) raa_.r[j] = std::max(kRaaEps, 0.1 * raa_.r[j]);
  subproblem(1, f0x, fx);
  solve_approx(xmma);
}

void DeviceGCMMAm::inner(double *xmma, double f0xnew, const double *fxnew, const double *xval,
                         double f0x, const double *df0dx, const double *fx, const double *dfdx,
                         const double *xmin, const double *xmax) {
  if (P_.iter == 0) throw Error(kNotReady, "GCMMA inner update before an outer update");
  load(xval, df0dx, dfdx, xmin, xmax);
  HIPCHK(hipMemcpyAsync(xm_.p, xmma, P_.n * sizeof(double), hipMemcpyHostToDevice, s_));
  // RaaUpdate (GCMMASolver.cpp:510-544)
  launch_gcmmam_raacof(P_, xm_, xv_, A_, part_, sums_, s_);
  HIPCHK(hipMemcpyAsync(hsum_.p, sums_.p, sizeof(double), hipMemcpyDeviceToHost, s_));
  HIPCHK(hipStreamSynchronize(s_));
  const double cof = std::max(1e-12, hsum_.p[0]);
  if (f0xnew > f0app_ + 0.5 * epsimin_)
    raa0_ = std::min(1.1 * (raa0_ + (1.0 / cof) * (f0xnew - f0app_)), 10.0 * raa0_);
  for (int j = 0; j < P_.m; ++j)
    if (fxnew[j] > fapp_.r[j] + 0.5 * epsimin_)
      raa_.r[j] = std::min(1.1 * (raa_.r[j] + (1.0 / cof) * (fxnew[j] - fapp_.r[j])),
                           10.0 * raa_.r[j]);
  subproblem(0, f0x, fx);
  solve_approx(xmma);
}

bool DeviceGCMMAm::concheck(double f0xnew, const double *fxnew) const {
  if (f0app_ + epsimin_ < f0xnew) return false;
  for (int j = 0; j < P_.m; ++j)
    if (fapp_.r[j] + epsimin_ < fxnew[j]) return false;
  return true;
}

void DeviceGCMMAm::state(double *out) const {
  const int m = P_.m;
  out[0] = raa0_;
  for (int j = 0; j < m; ++j) out[1 + j] = raa_.r[j];
  out[1 + m] = f0app_;
  for (int j = 0; j < m; ++j) out[2 + m + j] = fapp_.r[j];
}

}  // namespace icarat_hip
