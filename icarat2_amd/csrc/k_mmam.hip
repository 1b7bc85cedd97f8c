// k_mmam.hip -- the passes of MMA with m constraints (mmam.hpp), FP64.  Per
// Newton step of the dual interior-point iteration one element pass forms
// x(lambda) (XYZofLAMBDA, reference MMASolver.cpp:359-398) and, at that x, the
// m sums of DualGrad / DualResidual (:346-357, :184-212) and the m (m + 1) / 2
// sums of DualHess's PQ^T diag(df2) PQ (:238-344); one block then runs the
// scalar part: the Newton system (Factorize / Solve, :514-554), the search
// direction, DualLineSearch (:214-236) and the epsilon loop.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mmam.hpp"

namespace icarat_hip {

namespace {

constexpr int NT = 256;
constexpr int BLOCKS = 2048;
constexpr int MM = kMmaMaxM;
constexpr int NV = kMmamPassSums;

__device__ __forceinline__ int hidx(int j, int k) {  // (j <= k) in the packed upper triangle
  return MM + k * (k + 1) / 2 + j;
}

__global__ __launch_bounds__(NT) void k_mmam_gensub(MmamParams P, const double *__restrict__ x,
                                                    const double *__restrict__ dfdx,
                                                    const double *__restrict__ dgdx,
                                                    MmamArrays A, double *partials) {
  __shared__ double sh[MM * (NT / 64)];
  const long long n = P.n;
  const int m = P.m;
  double bs[MM];
#pragma unroll
  for (int j = 0; j < MM; ++j) bs[j] = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    const double xi = x[i], xo1 = A.xold1[i], xo2 = A.xold2[i];
    const double xmn = A.xmin[i], xmx = A.xmax[i];
    double lo, up;
    if (P.iter < 3) {
      lo = xi - P.asyminit * (xmx - xmn);
      up = xi + P.asyminit * (xmx - xmn);
    } else {
      const double osc = (xi - xo1) * (xo1 - xo2);
      const double gam = osc < 0.0 ? P.asymdec : (osc > 0.0 ? P.asyminc : 1.0);
      lo = xi - gam * (xo1 - A.low[i]);
      up = xi + gam * (A.upp[i] - xo1);
      const double w = fmax(P.xmamieps, xmx - xmn);
      lo = fmin(fmax(lo, xi - 100.0 * w), xi - 1.0e-5 * w);
      up = fmin(fmax(up, xi + 1.0e-5 * w), xi + 100.0 * w);
      const double xl = xmn - 1.0e-6, xh = xmx + 1.0e-6;
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
    double al = fmax(xmn, lo + P.albefa * (xi - lo));
    al = fmax(al, xi - P.move * (xmx - xmn));
    A.alpha[i] = fmin(al, xmx);
    double be = fmin(xmx, up - P.albefa * (up - xi));
    be = fmin(be, xi + P.move * (xmx - xmn));
    A.beta[i] = fmax(be, xmn);
    const double winv = 1.0 / fmax(P.xmamieps, xmx - xmn);
    const double ux = up - xi, xl = xi - lo;
    const double ux2 = ux * ux, xl2 = xl * xl;
    const double df = dfdx[i];
    const double r0 = 0.001 * fabs(df) + P.raa0 * winv;
    A.p0[i] = ux2 * (fmax(0.0, df) + r0);
    A.q0[i] = xl2 * (fmax(0.0, -df) + r0);
#pragma unroll
    for (int j = 0; j < MM; ++j) {
      if (j < m) {
        const double dg = dgdx[i * m + j];
        const double r1 = 0.001 * fabs(dg) + P.raa0 * winv;
        const double pij = ux2 * (fmax(0.0, dg) + r1), qij = xl2 * (fmax(0.0, -dg) + r1);
        A.P[j * n + i] = pij;
        A.Q[j * n + i] = qij;
        bs[j] += pij / ux + qij / xl;
      }
    }
    A.xold2[i] = xo1;
    A.xold1[i] = xi;
  }
  store_partials<NT, MM>(bs, partials, sh);
}

__global__ void k_mmam_init(MmamState *st, MmamState S0, const double *partials, unsigned G) {
  __shared__ double sh[MM * (NT / 64)];
  double tot[MM];
  sum_partials<NT, MM>(partials, G, tot, sh);
  if (threadIdx.x != 0) return;
  MmamState S = S0;
  for (int j = 0; j < S.m; ++j) S.b[j] += tot[j];  // b_j = -gx_j + sum (MMASolver.cpp:503-511)
  *st = S;
}

__global__ __launch_bounds__(NT) void k_mmam_pass(MmamParams P, const MmamState *st,
                                                  MmamArrays A, double *__restrict__ x,
                                                  double *partials) {
  __shared__ double sh[NV * (NT / 64)];
  if (st->done) return;
  const long long n = P.n;
  const int m = P.m;
  double lam[MM];
#pragma unroll
  for (int j = 0; j < MM; ++j) lam[j] = j < m ? fmax(st->lam[j], 0.0) : 0.0;
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    const double lo = A.low[i], up = A.upp[i];
    double pij[MM], qij[MM];
    double pl = A.p0[i], ql = A.q0[i];
#pragma unroll
    for (int j = 0; j < MM; ++j) {
      pij[j] = j < m ? A.P[j * n + i] : 0.0;
      qij[j] = j < m ? A.Q[j * n + i] : 0.0;
      if (j < m) {
        pl += pij[j] * lam[j];
        ql += qij[j] * lam[j];
      }
    }
    const double sp = sqrt(pl), sq = sqrt(ql);
    const double xp = (sp * lo + sq * up) / (sp + sq);
    double xi = xp < A.alpha[i] ? A.alpha[i] : xp;
    xi = xi > A.beta[i] ? A.beta[i] : xi;
    x[i] = xi;
    const double ux = up - xi, xl = xi - lo;
    const double ux2 = ux * ux, xl2 = xl * xl;
    double df2 = -1.0 / (2.0 * pl / (ux2 * ux) + 2.0 * ql / (xl2 * xl));
    if (xp < A.alpha[i] || xp > A.beta[i]) df2 = 0.0;
    double pq[MM];
#pragma unroll
    for (int j = 0; j < MM; ++j) {
      if (j < m) {
        v[j] += pij[j] / ux + qij[j] / xl;
        pq[j] = pij[j] / ux2 - qij[j] / xl2;
      }
    }
#pragma unroll
    for (int k = 0; k < MM; ++k)
#pragma unroll
      for (int j = 0; j <= k; ++j)
        if (k < m) v[hidx(j, k)] += pq[j] * df2 * pq[k];
  }
  store_partials<NT, NV>(v, partials, sh);
}

// Factorize / Solve of the reference (MMASolver.cpp:514-554): LU without
// pivoting in place, then the two triangular solves; x holds the right-hand
// side on entry
__device__ void lu_solve(double *K, double *x, int n) {
  for (int s = 0; s < n - 1; ++s)
    for (int i = s + 1; i < n; ++i) {
      K[i * n + s] = K[i * n + s] / K[s * n + s];
      for (int j = s + 1; j < n; ++j) K[i * n + j] = K[i * n + j] - K[i * n + s] * K[s * n + j];
    }
  for (int i = 1; i < n; ++i) {
    double a = 0.0;
    for (int j = 0; j < i; ++j) a = a - K[i * n + j] * x[j];
    x[i] = x[i] + a;
  }
  x[n - 1] = x[n - 1] / K[(n - 1) * n + (n - 1)];
  for (int i = n - 2; i >= 0; --i) {
    double a = x[i];
    for (int j = i + 1; j < n; ++j) a = a - K[i * n + j] * x[j];
    x[i] = a / K[i * n + i];
  }
}

// SolveDIP's scalar part in the host loop's order (as k_mma_newton for m = 1):
// phase 0 = the pass at the starting lambda, 1 = a pass after a Newton step
// (its sums give the residual first, then the next Newton system).
__global__ __launch_bounds__(NT) void k_mmam_newton(const double *partials, unsigned G,
                                                    MmamState *st) {
#pragma clang fp contract(off)
  __shared__ double sh[NV * (NT / 64)];
  __shared__ double hess[MM * MM];
  if (st->done) return;
  double tot[NV];
  sum_partials<NT, NV>(partials, G, tot, sh);
  if (threadIdx.x != 0) return;
  MmamState S = *st;
  const int m = S.m;
  // XYZofLAMBDA clamps lambda before the residual and the next step
  for (int j = 0; j < m; ++j)
    if (S.lam[j] < 0.0) S.lam[j] = 0.0;
  double y[MM], lamai = 0.0;
  for (int j = 0; j < m; ++j) {
    y[j] = fmax(0.0, S.lam[j] - S.c[j]);
    lamai += S.lam[j] * S.a[j];
  }
  const double z = fmax(0.0, 10.0 * (lamai - 1.0));
  if (S.phase == 1) {
    // DualResidual (MMASolver.cpp:184-212)
    double nr = 0.0;
    for (int j = 0; j < m; ++j) {
      const double r1 = -S.b[j] - S.a[j] * z - y[j] + S.mu[j] + tot[j];
      const double r2 = S.mu[j] * S.lam[j] - S.epsi;
      nr = fmax(nr, fabs(r1));
      nr = fmax(nr, fabs(r2));
    }
    S.err = nr;
  }
  for (;;) {
    if (S.phase == 1) {
      if (S.err > 0.9 * S.epsi && S.loop < 100) {
        ++S.loop;
        ++S.steps;
        const double epsi = S.epsi;
        double g[MM];
        for (int j = 0; j < m; ++j) {
          g[j] = -S.b[j] - S.a[j] * z - y[j] + tot[j];
          g[j] = -1.0 * g[j] - epsi / S.lam[j];
        }
        for (int j = 0; j < m; ++j)
          for (int k = 0; k < m; ++k) hess[j * m + k] = j <= k ? tot[hidx(j, k)] : tot[hidx(k, j)];
        for (int j = 0; j < m; ++j) {
          if (S.lam[j] > S.c[j]) hess[j * m + j] += -1.0;
          hess[j * m + j] += -S.mu[j] / S.lam[j];
        }
        if (lamai > 0.0)
          for (int j = 0; j < m; ++j)
            for (int k = 0; k < m; ++k) hess[j * m + k] += -10.0 * S.a[j] * S.a[k];
        double tr = 0.0;
        for (int j = 0; j < m; ++j) tr += hess[j * m + j];
        double corr = 1e-4 * tr / m;
        if (-1.0 * corr < 1.0e-7) corr = -1.0e-7;
        for (int j = 0; j < m; ++j) hess[j * m + j] += corr;
        double sd[2 * MM];
        if (m > 1) {
          lu_solve(hess, g, m);
          for (int j = 0; j < m; ++j) sd[j] = g[j];
        } else {
          sd[0] = g[0] / hess[0];
        }
        for (int j = 0; j < m; ++j)
          sd[m + j] = -S.mu[j] + epsi / S.lam[j] - sd[j] * S.mu[j] / S.lam[j];
        // DualLineSearch
        double th = 1.005;
        for (int j = 0; j < m; ++j) {
          if (th < -1.01 * sd[j] / S.lam[j]) th = -1.01 * sd[j] / S.lam[j];
          if (th < -1.01 * sd[m + j] / S.mu[j]) th = -1.01 * sd[m + j] / S.mu[j];
        }
        th = 1.0 / th;
        for (int j = 0; j < m; ++j) {
          S.lam[j] = S.lam[j] + th * sd[j];
          S.mu[j] = S.mu[j] + th * sd[m + j];
        }
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

// GCMMA (reference GCMMASolver.cpp): the subproblem at xval -- with asymp,
// Asymp's asymptotes first (:432-508) and the history shift of OuterUpdate
// (:66-90) -- as GenSub (:546-596): move limits without a second clamp to the
// bounds, raa0 in the objective and raa[j] in the constraints; partials: the
// objective's sum p0/(U-x) + q0/(x-L) and the m constraint sums.
constexpr int NG = 1 + MM;
__global__ __launch_bounds__(NT) void k_gcmmam_sub(MmamParams P, int asymp,
                                                   const double *__restrict__ x,
                                                   const double *__restrict__ df0dx,
                                                   const double *__restrict__ dfdx, double raa0,
                                                   GcmmaRaa raa, MmamArrays A, double *partials) {
  __shared__ double sh[NG * (NT / 64)];
  const long long n = P.n;
  const int m = P.m;
  double v[NG];
#pragma unroll
  for (int k = 0; k < NG; ++k) v[k] = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    const double xi = x[i];
    const double xmn = A.xmin[i], xmx = A.xmax[i];
    double lo, up;
    if (asymp) {
      const double xo1 = A.xold1[i], xo2 = A.xold2[i];
      if (P.iter < 3) {
        lo = xi - P.asyminit * (xmx - xmn);
        up = xi + P.asyminit * (xmx - xmn);
      } else {
        const double osc = (xi - xo1) * (xo1 - xo2);
        const double gam = osc < 0.0 ? P.asymdec : (osc > 0.0 ? P.asyminc : 1.0);
        lo = xi - gam * (xo1 - A.low[i]);
        up = xi + gam * (A.upp[i] - xo1);
        const double w = fmax(P.xmamieps, xmx - xmn);
        lo = fmin(fmax(lo, xi - 100.0 * w), xi - 1e-5 * w);
        up = fmin(fmax(up, xi + 1e-5 * w), xi + 100.0 * w);
        const double xl = xmn - 1.0e-6, xh = xmx + 1.0e-6;
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
    } else {
      lo = A.low[i];
      up = A.upp[i];
    }
    A.alpha[i] = fmax(fmax(xmn, lo + P.albefa * (xi - lo)), xi - P.move * (xmx - xmn));
    A.beta[i] = fmin(fmin(xmx, up - P.albefa * (up - xi)), xi + P.move * (xmx - xmn));
    const double uxinv = 1.0 / (up - xi), xlinv = 1.0 / (xi - lo);
    const double ux2 = (up - xi) * (up - xi), xl2 = (xi - lo) * (xi - lo);
    const double winv = 1.0 / fmax(P.xmamieps, xmx - xmn);
    const double df = df0dx[i];
    const double r0 = 0.001 * fabs(df) + raa0 * winv;
    const double p0 = ux2 * (fmax(0.0, df) + r0), q0 = xl2 * (fmax(0.0, -df) + r0);
    A.p0[i] = p0;
    A.q0[i] = q0;
    v[0] += p0 * uxinv + q0 * xlinv;
#pragma unroll
    for (int j = 0; j < MM; ++j) {
      if (j < m) {
        const double dg = dfdx[i * m + j];
        const double r1 = 0.001 * fabs(dg) + raa.r[j] * winv;
        const double pij = ux2 * (fmax(0.0, dg) + r1), qij = xl2 * (fmax(0.0, -dg) + r1);
        A.P[j * n + i] = pij;
        A.Q[j * n + i] = qij;
        v[1 + j] += pij * uxinv + qij * xlinv;
      }
    }
  }
  store_partials<NT, NG>(v, partials, sh);
}

// ComputeApprox's sums at xmma (GCMMASolver.cpp:598-620): the objective's and
// the m constraints' p/(U-x) + q/(x-L)
__global__ __launch_bounds__(NT) void k_gcmmam_approx(MmamParams P, const double *__restrict__ xm,
                                                      MmamArrays A, double *partials) {
  __shared__ double sh[NG * (NT / 64)];
  const long long n = P.n;
  const int m = P.m;
  double v[NG];
#pragma unroll
  for (int k = 0; k < NG; ++k) v[k] = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    const double uxinv = 1.0 / (A.upp[i] - xm[i]), xlinv = 1.0 / (xm[i] - A.low[i]);
    v[0] += A.p0[i] * uxinv + A.q0[i] * xlinv;
#pragma unroll
    for (int j = 0; j < MM; ++j)
      if (j < m) v[1 + j] += A.P[j * n + i] * uxinv + A.Q[j * n + i] * xlinv;
  }
  store_partials<NT, NG>(v, partials, sh);
}

// RaaUpdate's raacof (GCMMASolver.cpp:510-526): sum ((xm-xv)/(U-xm))
// ((xm-xv)/(xm-L)) (U-L)/xmami
__global__ __launch_bounds__(NT) void k_gcmmam_raacof(MmamParams P, const double *__restrict__ xm,
                                                      const double *__restrict__ xv,
                                                      MmamArrays A, double *partials) {
  __shared__ double sh[NT / 64];
  const long long n = P.n;
  double acc = 0.0;
  const long long stride = (long long)gridDim.x * NT;
  for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += stride) {
    const double xmami = fmax(P.xmamieps, A.xmax[i] - A.xmin[i]);
    const double lo = A.low[i], up = A.upp[i];
    const double xxux = (xm[i] - xv[i]) / (up - xm[i]);
    const double xxxl = (xm[i] - xv[i]) / (xm[i] - lo);
    acc += (xxux * xxxl) * ((up - lo) / xmami);
  }
  double v[1] = {acc};
  store_partials<NT, 1>(v, partials, sh);
}

// the blocks' partials summed in a fixed order into out[0 .. nv)
template <int NVS>
__global__ void k_mmam_sum(const double *partials, unsigned G, int nv, double *out) {
  __shared__ double sh[NVS * (NT / 64)];
  double tot[NVS];
  sum_partials<NT, NVS>(partials, G, tot, sh);
  if (threadIdx.x == 0)
    for (int k = 0; k < nv; ++k) out[k] = tot[k];
}

}  // namespace

unsigned mmam_blocks(long long n) {
  return (unsigned)std::max<long long>(1, std::min<long long>(BLOCKS, (n + NT - 1) / NT));
}

void launch_mmam_gensub(const MmamParams &P, const double *x, const double *dfdx,
                        const double *dgdx, const MmamArrays &A, double *partials, hipStream_t s) {
  hipLaunchKernelGGL(k_mmam_gensub, dim3(mmam_blocks(P.n)), dim3(NT), 0, s, P, x, dfdx, dgdx, A,
                     partials);
  HIPCHK(hipGetLastError());
}

void launch_mmam_init(const MmamParams &P, const double *gx_host, double a, double c,
                      double epsimin, const double *partials, MmamState *st, hipStream_t s) {
  MmamState S{};
  S.m = P.m;
  for (int j = 0; j < P.m; ++j) {
    S.b[j] = -gx_host[j];
    S.a[j] = a;
    S.c[j] = c;
    S.lam[j] = c / 2.0;
    S.mu[j] = 1.0;
  }
  S.epsi = 1.0;
  S.err = 1.0;
  S.epsimin = epsimin;
  hipLaunchKernelGGL(k_mmam_init, dim3(1), dim3(NT), 0, s, st, S, partials, mmam_blocks(P.n));
  HIPCHK(hipGetLastError());
}

void launch_mmam_step(const MmamParams &P, const MmamArrays &A, double *x, double *partials,
                      MmamState *st, hipStream_t s) {
  const unsigned G = mmam_blocks(P.n);
  hipLaunchKernelGGL(k_mmam_pass, dim3(G), dim3(NT), 0, s, P, st, A, x, partials);
  hipLaunchKernelGGL(k_mmam_newton, dim3(1), dim3(NT), 0, s, partials, G, st);
  HIPCHK(hipGetLastError());
}

}  // namespace icarat_hip

namespace icarat_hip {

void launch_gcmmam_sub(const MmamParams &P, int asymp, const double *x, const double *df0dx,
                       const double *dfdx, double raa0, const GcmmaRaa &raa, const MmamArrays &A,
                       double *partials, double *out, hipStream_t s) {
  const unsigned G = mmam_blocks(P.n);
  hipLaunchKernelGGL(k_gcmmam_sub, dim3(G), dim3(NT), 0, s, P, asymp, x, df0dx, dfdx, raa0, raa,
                     A, partials);
  hipLaunchKernelGGL(k_mmam_sum<NG>, dim3(1), dim3(NT), 0, s, partials, G, 1 + P.m, out);
  HIPCHK(hipGetLastError());
}

void launch_gcmmam_approx(const MmamParams &P, const double *xm, const MmamArrays &A,
                          double *partials, double *out, hipStream_t s) {
  const unsigned G = mmam_blocks(P.n);
  hipLaunchKernelGGL(k_gcmmam_approx, dim3(G), dim3(NT), 0, s, P, xm, A, partials);
  hipLaunchKernelGGL(k_mmam_sum<NG>, dim3(1), dim3(NT), 0, s, partials, G, 1 + P.m, out);
  HIPCHK(hipGetLastError());
}

void launch_gcmmam_raacof(const MmamParams &P, const double *xm, const double *xv,
                          const MmamArrays &A, double *partials, double *out, hipStream_t s) {
  const unsigned G = mmam_blocks(P.n);
  hipLaunchKernelGGL(k_gcmmam_raacof, dim3(G), dim3(NT), 0, s, P, xm, xv, A, partials);
  hipLaunchKernelGGL(k_mmam_sum<1>, dim3(1), dim3(NT), 0, s, partials, G, 1, out);
  HIPCHK(hipGetLastError());
}


}  // namespace icarat_hip
