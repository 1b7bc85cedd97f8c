// mma_kernels.hpp -- launchers of the element-parallel MMA passes (k_mma.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "hip_util.hpp"

namespace icarat_hip {

// Parameters of MMASolver with the reference defaults
// (optimization/opt-solver/MMASolver.cpp:41-56).
struct MmaParams {
  int iter;
  double xmin, xmax;
  double xmamieps = 1.0e-5, raa0 = 1.0e-5;
  double move = 0.15, albefa = 0.2, asyminit = 0.2, asymdec = 0.7, asyminc = 1.2;
};

struct MmaArrays {
  double *low, *upp, *alpha, *beta, *p0, *q0, *p1, *q1, *xold1, *xold2;
};

// Builds the subproblem at x; out[0] = sum p1/(U-x) + q1/(x-L).
void launch_mma_gensub(long long n, const double *x, const double *dfdx, const double *dgdx,
                       const MmaParams &P, const MmaArrays &A, double *partials,
                       unsigned *counter, double *out, hipStream_t s);
// x = x(lambda); out[0] = Sg, out[1] = Sh at that x.
void launch_mma_dual(long long n, double lam, const MmaArrays &A, double *x, double *partials,
                     unsigned *counter, double *out, hipStream_t s);

// State of the dual interior-point iteration for one constraint, kept in
// device memory so that the Newton steps run without host round trips.
struct MmaDualState {
  double lam, mu, epsi, err, b, a, c, epsimin;
  int outer, loop, steps, phase, done, pad;
};
// st = the starting point lambda = c/2, mu = 1, epsi = err = 1 with
// b = b0 (+ bsum[0] when bsum is given: the subproblem's sum on the device)
void launch_mma_dual_init(MmaDualState *st, double b0, const double *bsum, double a, double c,
                          double epsimin, hipStream_t s);
// One evaluation x(lambda), Sg, Sh at the state's lambda, then the scalar
// Newton logic up to the next lambda that needs an evaluation (or done).
// Both kernels return at once when the state is done.
void launch_mma_dual_step(long long n, const MmaArrays &A, double *x, double *partials,
                          MmaDualState *st, hipStream_t s);

// GCMMA (reference optimization/opt-solver/GCMMASolver.cpp) for one constraint:
// the subproblem at x (asymp != 0: move the asymptotes and shift the history
// first, as OuterUpdate does); out[0] = sum p0/(U-x) + q0/(x-L), out[1] = the
// same of p1, q1.
void launch_gcmma_sub(long long n, const double *x, const double *df0, const double *dg,
                      const MmaParams &P, double raa0, double raa1, int asymp, const MmaArrays &A,
                      double *partials, double *out, hipStream_t s);
// out[0..1] = the two approximation sums at xm
void launch_gcmma_approx(long long n, const double *xm, const MmaArrays &A, double *partials,
                         double *out, hipStream_t s);
// out[0] = sum ((xm-xv)/(U-xm)) ((xm-xv)/(xm-L)) (U-L) winv
void launch_gcmma_raacof(long long n, const double *xm, const double *xv, double winv,
                         const MmaArrays &A, double *partials, double *out, hipStream_t s);

}  // namespace icarat_hip
