// mmam.hpp -- Svanberg's MMA and GCMMA with m constraints (1 <= m <=
// kMmaMaxM) on the device: MMASolver(n, m, a, c, d) and GCMMASolver(n, m, a,
// c, d) of the reference (optimization/opt-solver/MMASolver.hpp:62-116,
// MMASolver.cpp, GCMMASolver.hpp:62-85, GCMMASolver.cpp), FP64.  The
// subproblem, x(lambda) and the dual sums are element-parallel passes; the
// m x m Newton system of the dual interior-point iteration, its line search
// and the epsilon loop run in one device block, so an update needs no host
// round trip per Newton step.
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "hip_util.hpp"

namespace icarat_hip {

constexpr int kMmaMaxM = 8;

struct MmamParams {
  long long n;
  int m, iter;
  double xmamieps = 1.0e-5, raa0 = 1.0e-5;
  double move = 0.15, albefa = 0.2, asyminit = 0.2, asymdec = 0.7, asyminc = 1.2;
};

// element arrays (length n; P and Q constraint-major, m x n: pij at j n + i)
struct MmamArrays {
  double *low, *upp, *alpha, *beta, *p0, *q0, *P, *Q, *xold1, *xold2, *xmin, *xmax;
};

// the dual iteration's state (SolveDIP, MMASolver.cpp:88-154)
struct MmamState {
  double lam[kMmaMaxM], mu[kMmaMaxM], b[kMmaMaxM], a[kMmaMaxM], c[kMmaMaxM];
  double epsi, err, epsimin;
  int m, outer, loop, steps, phase, done;
};

// the partials of one pass per quantity (blocks of the element passes)
unsigned mmam_blocks(long long n);
// quantities per pass: m gradient sums and m (m + 1) / 2 Hessian sums
constexpr int kMmamPassSums = kMmaMaxM + kMmaMaxM * (kMmaMaxM + 1) / 2;

// GenSub (MMASolver.cpp:400-512): asymptotes, bounds, the approximation, the
// history shift; partials of sum_i pij/(U-x) + qij/(x-L) per constraint.
// dgdx in the reference's layout (dgdx[i m + j]).
void launch_mmam_gensub(const MmamParams &P, const double *x, const double *dfdx,
                        const double *dgdx, const MmamArrays &A, double *partials, hipStream_t s);
// the starting state: b_j = -gx_j + the gensub sums, lambda = c/2, mu = 1
void launch_mmam_init(const MmamParams &P, const double *gx_host, double a, double c,
                      double epsimin, const double *partials, MmamState *st, hipStream_t s);
// one pass x(lambda) with its dual sums, then the Newton logic up to the next
// lambda that needs a pass (or done); both return at once once done
void launch_mmam_step(const MmamParams &P, const MmamArrays &A, double *x, double *partials,
                      MmamState *st, hipStream_t s);

// GCMMA (reference GCMMASolver.cpp) passes: the subproblem at x (asymp != 0:
// Asymp's asymptotes and the history shift first) with out[0] = the
// objective's sum p0/(U-x) + q0/(x-L) and out[1 + j] the constraints' sums;
// the approximation sums at xm (same layout); raacof at xm, xv (out[0]).
// out is device memory.
struct GcmmaRaa {
  double r[kMmaMaxM];
};
void launch_gcmmam_sub(const MmamParams &P, int asymp, const double *x, const double *df0dx,
                       const double *dfdx, double raa0, const GcmmaRaa &raa, const MmamArrays &A,
                       double *partials, double *out, hipStream_t s);
void launch_gcmmam_approx(const MmamParams &P, const double *xm, const MmamArrays &A,
                          double *partials, double *out, hipStream_t s);
void launch_gcmmam_raacof(const MmamParams &P, const double *xm, const double *xv,
                          const MmamArrays &A, double *partials, double *out, hipStream_t s);

// MMASolver(n, m, a, c, d) with its own stream; host pointers in update.
class DeviceMMAm {
 public:
  DeviceMMAm(long long n, int m, double a, double c, double d);
  ~DeviceMMAm();
  DeviceMMAm(const DeviceMMAm &) = delete;
  DeviceMMAm &operator=(const DeviceMMAm &) = delete;
  // SetMoveLimit, SetOscillation, SetAsymptotes (MMASolver.hpp:68-74)
  void set_params(double move, double albefa, double asyminit, double asymdec, double asyminc);
  void reset() { P_.iter = 0; }  // Reset (MMASolver.hpp:80)
  // Update(xval, dfdx, gx, dgdx, xmin, xmax) (MMASolver.cpp:66-82), host
  // arrays: x (n, in/out), dfdx (n), gx (m), dgdx (n m, dgdx[i m + j]), xmin
  // and xmax (n).  Returns the interior-point outer loops.
  int update(double *x, const double *dfdx, const double *gx, const double *dgdx,
             const double *xmin, const double *xmax);
  // the asymptotes of the last update (host arrays of n)
  void asymptotes(double *low, double *upp);
  int last_newton_steps() const { return steps_; }

 private:
  hipStream_t s_ = nullptr;
  MmamParams P_;
  double a_, c_, d_;
  int steps_ = 0;
  DevBuf<double> x_, df_, dg_, part_;
  DevBuf<double> low_, upp_, alpha_, beta_, p0_, q0_, pij_, qij_, xold1_, xold2_, xmin_, xmax_;
  DevBuf<MmamState> st_;
  PinnedBuf<MmamState> hst_;
  MmamArrays A_;
};

// GCMMASolver(n, m, a, c, d) (reference GCMMASolver.hpp:62-85) with its own
// stream; host pointers.  The scalars (raa0, raa, the approximation
// constants, ConCheck) stay on the host; the element passes and the dual
// iteration run on the device.
class DeviceGCMMAm {
 public:
  DeviceGCMMAm(long long n, int m, double a, double c, double d);
  ~DeviceGCMMAm();
  DeviceGCMMAm(const DeviceGCMMAm &) = delete;
  DeviceGCMMAm &operator=(const DeviceGCMMAm &) = delete;
  void set_params(double move, double albefa, double asyminit, double asymdec, double asyminc);
  void reset() { P_.iter = 0; }  // Reset (GCMMASolver.hpp:83): the outer counter only
  // OuterUpdate / InnerUpdate (GCMMASolver.cpp:66-113): xmma (n) receives the
  // new design (InnerUpdate reads it first); fx, fxnew (m); dfdx (n m,
  // dfdx[i m + j]); xval, df0dx, xmin, xmax (n)
  void outer(double *xmma, const double *xval, double f0x, const double *df0dx, const double *fx,
             const double *dfdx, const double *xmin, const double *xmax);
  void inner(double *xmma, double f0xnew, const double *fxnew, const double *xval, double f0x,
             const double *df0dx, const double *fx, const double *dfdx, const double *xmin,
             const double *xmax);
  bool concheck(double f0xnew, const double *fxnew) const;  // ConCheck (:115-133)
  // {raa0, raa[m], f0app, fapp[m]}
  void state(double *out) const;

 private:
  void load(const double *xval, const double *df0dx, const double *dfdx, const double *xmin,
            const double *xmax);
  void subproblem(int asymp, double f0x, const double *fx);
  void solve_approx(double *xmma);
  hipStream_t s_ = nullptr;
  MmamParams P_;
  double a_, c_, d_, epsimin_;
  double raa0_ = 1.0, r0_ = 0.0, f0app_ = 0.0;
  GcmmaRaa raa_, r_, fapp_;
  DevBuf<double> xv_, xm_, df_, dg_, part_, sums_;
  DevBuf<double> low_, upp_, alpha_, beta_, p0_, q0_, pij_, qij_, xold1_, xold2_, xmin_, xmax_;
  DevBuf<MmamState> st_;
  PinnedBuf<MmamState> hst_;
  PinnedBuf<double> hsum_;
  MmamArrays A_;
};

}  // namespace icarat_hip
