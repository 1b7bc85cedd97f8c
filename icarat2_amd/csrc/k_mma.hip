// k_mma.hip -- element-parallel passes of Svanberg's MMA with one constraint
// (the SIMP volume constraint), FP64.  The dual interior-point iteration of
// the reference (optimization/opt-solver/MMASolver.cpp:88-154) only needs, per
// Newton step, x(lambda) and two sums over the elements:
//   Sg = sum p1/(U-x) + q1/(x-L)                      (DualGrad, DualResidual)
//   Sh = sum PQ^2 df2, PQ = p1/(U-x)^2 - q1/(x-L)^2   (DualHess)
// and one pass produces all three.
#include <hip/hip_runtime.h>

#include "mma_kernels.hpp"

namespace icarat_hip {

namespace {

constexpr int NT = 256;
constexpr int BLOCKS = 2048;

// Asymptotes, move limits and the convex separable approximation
// (reference MMASolver.cpp:400-512); also shifts the iterate history.
__global__ __launch_bounds__(NT) void k_mma_gensub(long long n, const double *__restrict__ x,
                                                   const double *__restrict__ dfdx,
                                                   const double *__restrict__ dgdx, MmaParams P,
                                                   MmaArrays A, double *partials,
                                                   unsigned *counter, double *out) {
  __shared__ double sh[4];
  const double span = P.xmax - P.xmin;
  const double w = fmax(P.xmamieps, span);
  const double winv = 1.0 / w;
  double bsum = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    const double xi = x[i];
    const double xo1 = A.xold1[i], xo2 = A.xold2[i];
    double lo, up;
    if (P.iter < 3) {
      lo = xi - P.asyminit * span;
      up = xi + P.asyminit * span;
    } else {
      const double osc = (xi - xo1) * (xo1 - xo2);
      const double gam = osc < 0.0 ? P.asymdec : (osc > 0.0 ? P.asyminc : 1.0);
      lo = xi - gam * (xo1 - A.low[i]);
      up = xi + gam * (A.upp[i] - xo1);
      lo = fmin(fmax(lo, xi - 100.0 * w), xi - 1.0e-5 * w);
      up = fmin(fmax(up, xi + 1.0e-5 * w), xi + 100.0 * w);
      const double xl = P.xmin - 1.0e-6, xh = P.xmax + 1.0e-6;
      if (xi < xl) {
        lo = xi - (xh - xi) / 0.9;
        up = xi + (xh - xi) / 0.9;
      }
      if (xi > xh) {
        lo = xi - (xi - xl) / 0.9;
        up = xi + (xi - xl) / 0.9;
      }
    }
    A.low[i] = lo;
    A.upp[i] = up;
    double al = fmax(P.xmin, lo + P.albefa * (xi - lo));
    al = fmax(al, xi - P.move * span);
    A.alpha[i] = fmin(al, P.xmax);
    double be = fmin(P.xmax, up - P.albefa * (up - xi));
    be = fmin(be, xi + P.move * span);
    A.beta[i] = fmax(be, P.xmin);
    const double ux = up - xi, xl2 = xi - lo;
    const double ux2 = ux * ux, xl2sq = xl2 * xl2;
    const double df = dfdx[i], dg = dgdx[i];
    const double r0 = 0.001 * fabs(df) + P.raa0 * winv;
    const double r1 = 0.001 * fabs(dg) + P.raa0 * winv;
    const double p1 = ux2 * (fmax(0.0, dg) + r1), q1 = xl2sq * (fmax(0.0, -dg) + r1);
    A.p0[i] = ux2 * (fmax(0.0, df) + r0);
    A.q0[i] = xl2sq * (fmax(0.0, -df) + r0);
    A.p1[i] = p1;
    A.q1[i] = q1;
    A.xold2[i] = xo1;
    A.xold1[i] = xi;
    bsum += p1 / ux + q1 / xl2;
  }
  double v[1] = {bsum};
  store_partials<NT, 1>(v, partials, sh);
}

// x(lambda) (XYZofLAMBDA, reference MMASolver.cpp:359-398) and the two dual
// sums at that x (DualGrad :346-357, DualHess :238-344).
__global__ __launch_bounds__(NT) void k_mma_dual(long long n, double lam, MmaArrays A,
                                                 double *__restrict__ x, double *partials,
                                                 unsigned *counter, double *out) {
  __shared__ double sh[8];
  double sg = 0.0, shs = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    const double lo = A.low[i], up = A.upp[i];
    const double pl = A.p0[i] + A.p1[i] * lam;
    const double ql = A.q0[i] + A.q1[i] * lam;
    const double sp = sqrt(pl), sq = sqrt(ql);
    const double xp = (sp * lo + sq * up) / (sp + sq);
    double xi = xp < A.alpha[i] ? A.alpha[i] : xp;
    xi = xi > A.beta[i] ? A.beta[i] : xi;
    x[i] = xi;
    const double ux = up - xi, xl = xi - lo;
    sg += A.p1[i] / ux + A.q1[i] / xl;
    const double pq = A.p1[i] / (ux * ux) - A.q1[i] / (xl * xl);
    double df2 = -1.0 / (2.0 * pl / (ux * ux * ux) + 2.0 * ql / (xl * xl * xl));
    if (xp < A.alpha[i] || xp > A.beta[i]) df2 = 0.0;
    shs += pq * df2 * pq;
  }
  double v[2] = {sg, shs};
  store_partials<NT, 2>(v, partials, sh);
}

}  // namespace

void launch_mma_gensub(long long n, const double *x, const double *dfdx, const double *dgdx,
                       const MmaParams &P, const MmaArrays &A, double *partials,
                       unsigned *counter, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_mma_gensub, dim3(BLOCKS), dim3(NT), 0, s, n, x, dfdx, dgdx, P, A,
                     partials, counter, out);
  HIPCHK(hipGetLastError());
  launch_finalize(kFinOut, 1, partials, BLOCKS, nullptr, nullptr, out, s);
}

void launch_mma_dual(long long n, double lam, const MmaArrays &A, double *x, double *partials,
                     unsigned *counter, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_mma_dual, dim3(BLOCKS), dim3(NT), 0, s, n, lam, A, x, partials, counter,
                     out);
  HIPCHK(hipGetLastError());
  launch_finalize(kFinOut, 2, partials, BLOCKS, nullptr, nullptr, out, s);
}

}  // namespace icarat_hip
