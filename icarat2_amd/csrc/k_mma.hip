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

// k_mma_dual at the lambda of a device state; no work once the state is done
__global__ __launch_bounds__(NT) void k_mma_dual_dev(long long n, const MmaDualState *st,
                                                     MmaArrays A, double *__restrict__ x,
                                                     double *partials) {
  __shared__ double sh[8];
  if (st->done) return;
  const double lam = st->lam;
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

__device__ __forceinline__ double dmax(double a, double b) { return a < b ? b : a; }

__global__ void k_mma_dual_init(MmaDualState *st, double b0, const double *bsum, double a,
                                double c, double epsimin) {
#pragma clang fp contract(off)
  MmaDualState S{};
  S.b = bsum ? b0 + bsum[0] : b0;
  S.a = a;
  S.c = c;
  S.epsimin = epsimin;
  S.lam = c / 2.0;
  S.mu = 1.0;
  S.epsi = 1.0;
  S.err = 1.0;
  *st = S;
}

// SolveDIP's scalar part (reference MMASolver.cpp:88-154 with the Newton step
// of :156-236, GCMMASolver.cpp:128-196), m = 1, in the host loop's order:
//   pass(lam); while (epsi > epsimin && outer < 100) { ++outer; loop = 0;
//     while (err > 0.9 epsi && loop < 100) { Newton step; pass(lam); err }
//     epsi *= 0.1 }
// Each launch consumes one pass: phase 0 = the first pass, 1 = a pass after a
// Newton step (err is updated from it).
__global__ __launch_bounds__(NT) void k_mma_newton(const double *partials, unsigned G,
                                                   MmaDualState *st) {
#pragma clang fp contract(off)
  __shared__ double sh[8];
  if (st->done) return;
  double tot[2];
  sum_partials<NT, 2>(partials, G, tot, sh);
  if (threadIdx.x != 0) return;
  MmaDualState S = *st;
  const double Sg = tot[0], Sh = tot[1];
  const double a = S.a, c = S.c, b = S.b;
  if (S.phase == 1) {
    const double y = dmax(0.0, S.lam - c);
    const double z = dmax(0.0, 10.0 * (S.lam * a - 1.0));
    const double r1 = -b - a * z - y + S.mu + Sg;
    const double r2 = S.mu * S.lam - S.epsi;
    S.err = dmax(fabs(r1), fabs(r2));
  }
  for (;;) {
    if (S.phase == 1) {
      if (S.err > 0.9 * S.epsi && S.loop < 100) {
        ++S.loop;
        ++S.steps;
        double lam = dmax(S.lam, 0.0);
        const double mu = S.mu, epsi = S.epsi;
        const double y = dmax(0.0, lam - c);
        const double z = dmax(0.0, 10.0 * (lam * a - 1.0));
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
        double th = 1.005;
        th = dmax(th, -1.01 * s0 / lam);
        th = dmax(th, -1.01 * s1 / mu);
        th = 1.0 / th;
        S.lam = lam + th * s0;
        S.mu = mu + th * s1;
        break;  // the next launch evaluates x(lambda)
      }
      S.epsi *= 0.1;
      S.phase = 2;
    }
    if (S.epsi > S.epsimin && S.outer < 100) {
      ++S.outer;
      S.loop = 0;
      S.phase = 1;
      continue;
    }
    S.done = 1;
    break;
  }
  *st = S;
}

// GCMMA subproblem (reference GCMMASolver.cpp:486-560 Asymp when asymp != 0,
// then :593-640 GenSub): MMA's asymptotes and history shift, move limits
// without a second clamp to the bounds, raa0 in the objective and raa1 in the
// constraint.  Partials: sum p0/(U-x) + q0/(x-L) and sum p1/(U-x) + q1/(x-L).
__global__ __launch_bounds__(NT) void k_gcmma_sub(long long n, const double *__restrict__ x,
                                                  const double *__restrict__ df0,
                                                  const double *__restrict__ dg, MmaParams P,
                                                  double raa0, double raa1, int asymp,
                                                  MmaArrays A, double *partials) {
  __shared__ double sh[8];
  const double span = P.xmax - P.xmin;
  const double w = fmax(P.xmamieps, span);
  const double winv = 1.0 / w;
  double s0 = 0.0, s1 = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    const double xi = x[i];
    double lo = A.low[i], up = A.upp[i];
    if (asymp) {
      const double xo1 = A.xold1[i], xo2 = A.xold2[i];
      if (P.iter < 3) {
        lo = xi - P.asyminit * span;
        up = xi + P.asyminit * span;
      } else {
        const double osc = (xi - xo1) * (xo1 - xo2);
        const double gam = osc < 0.0 ? P.asymdec : (osc > 0.0 ? P.asyminc : 1.0);
        lo = xi - gam * (xo1 - lo);
        up = xi + gam * (up - xo1);
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
      A.xold2[i] = xo1;
      A.xold1[i] = xi;
    }
    A.alpha[i] = fmax(fmax(P.xmin, lo + P.albefa * (xi - lo)), xi - P.move * span);
    A.beta[i] = fmin(fmin(P.xmax, up - P.albefa * (up - xi)), xi + P.move * span);
    const double ux = up - xi, xl2 = xi - lo;
    const double uxinv = 1.0 / ux, xlinv = 1.0 / xl2;
    const double ux2 = ux * ux, xl2sq = xl2 * xl2;
    const double f = df0[i], g = dg[i];
    const double pq0 = 0.001 * fabs(f) + raa0 * winv;
    const double pq1 = 0.001 * fabs(g) + raa1 * winv;
    const double p0 = ux2 * (fmax(0.0, f) + pq0), q0 = xl2sq * (fmax(0.0, -f) + pq0);
    const double p1 = ux2 * (fmax(0.0, g) + pq1), q1 = xl2sq * (fmax(0.0, -g) + pq1);
    A.p0[i] = p0;
    A.q0[i] = q0;
    A.p1[i] = p1;
    A.q1[i] = q1;
    s0 += p0 * uxinv + q0 * xlinv;
    s1 += p1 * uxinv + q1 * xlinv;
  }
  double v[2] = {s0, s1};
  store_partials<NT, 2>(v, partials, sh);
}

// ComputeApprox (GCMMASolver.cpp:642-661) without the constants: the two
// approximation sums at xm
__global__ __launch_bounds__(NT) void k_gcmma_approx(long long n, const double *__restrict__ xm,
                                                     MmaArrays A, double *partials) {
  __shared__ double sh[8];
  double s0 = 0.0, s1 = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    const double uxinv = 1.0 / (A.upp[i] - xm[i]), xlinv = 1.0 / (xm[i] - A.low[i]);
    s0 += A.p0[i] * uxinv + A.q0[i] * xlinv;
    s1 += A.p1[i] * uxinv + A.q1[i] * xlinv;
  }
  double v[2] = {s0, s1};
  store_partials<NT, 2>(v, partials, sh);
}

// the raa growth coefficient of RaaUpdate (GCMMASolver.cpp:562-576)
__global__ __launch_bounds__(NT) void k_gcmma_raacof(long long n, const double *__restrict__ xm,
                                                     const double *__restrict__ xv, double winv,
                                                     MmaArrays A, double *partials) {
  __shared__ double sh[4];
  double s = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    const double up = A.upp[i], lo = A.low[i], d = xm[i] - xv[i];
    s += (d / (up - xm[i])) * (d / (xm[i] - lo)) * ((up - lo) * winv);
  }
  double v[1] = {s};
  store_partials<NT, 1>(v, partials, sh);
}

}  // namespace

void launch_gcmma_sub(long long n, const double *x, const double *df0, const double *dg,
                      const MmaParams &P, double raa0, double raa1, int asymp, const MmaArrays &A,
                      double *partials, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_gcmma_sub, dim3(BLOCKS), dim3(NT), 0, s, n, x, df0, dg, P, raa0, raa1,
                     asymp, A, partials);
  HIPCHK(hipGetLastError());
  launch_finalize(kFinOut, 2, partials, BLOCKS, nullptr, nullptr, out, s);
}

void launch_gcmma_approx(long long n, const double *xm, const MmaArrays &A, double *partials,
                         double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_gcmma_approx, dim3(BLOCKS), dim3(NT), 0, s, n, xm, A, partials);
  HIPCHK(hipGetLastError());
  launch_finalize(kFinOut, 2, partials, BLOCKS, nullptr, nullptr, out, s);
}

void launch_gcmma_raacof(long long n, const double *xm, const double *xv, double winv,
                         const MmaArrays &A, double *partials, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_gcmma_raacof, dim3(BLOCKS), dim3(NT), 0, s, n, xm, xv, winv, A, partials);
  HIPCHK(hipGetLastError());
  launch_finalize(kFinOut, 1, partials, BLOCKS, nullptr, nullptr, out, s);
}

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

void launch_mma_dual_init(MmaDualState *st, double b0, const double *bsum, double a, double c,
                          double epsimin, hipStream_t s) {
  hipLaunchKernelGGL(k_mma_dual_init, dim3(1), dim3(1), 0, s, st, b0, bsum, a, c, epsimin);
  HIPCHK(hipGetLastError());
}

void launch_mma_dual_step(long long n, const MmaArrays &A, double *x, double *partials,
                          MmaDualState *st, hipStream_t s) {
  hipLaunchKernelGGL(k_mma_dual_dev, dim3(BLOCKS), dim3(NT), 0, s, n, st, A, x, partials);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_mma_newton, dim3(1), dim3(NT), 0, s, partials, (unsigned)BLOCKS, st);
  HIPCHK(hipGetLastError());
}

}  // namespace icarat_hip
