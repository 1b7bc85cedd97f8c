// k_reduce.hip -- the single-block finalize kernel of every reduction: sums
// the producer blocks' partials in block order and runs the scalar step that
// consumes them (PCG alpha/beta and stop test, OC bisection, plain output).
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "hip_util.hpp"

namespace icarat_hip {

namespace {

constexpr int NT = 256;

template <int NV>
__global__ __launch_bounds__(NT) void k_finalize(int op, const double *__restrict__ partials,
                                                 unsigned G, PcgScalars *S, OcState *O,
                                                 double *out) {
  __shared__ double sh[4 * NV];
  if ((op == kFinPcgAlpha || op == kFinPcgBeta) && S->done) return;
  if (op == kFinOc && O->done) return;
  double tot[NV];
  sum_partials<NT, NV>(partials, G, tot, sh);
  if (threadIdx.x != 0) return;
  switch (op) {
    case kFinOut:
#pragma unroll
      for (int k = 0; k < NV; ++k) out[k] = tot[k];
      break;
    case kFinPcgInit:
      S->rho = tot[0];
      S->bnorm2 = tot[NV - 1];
      S->rr = tot[NV - 1];
      S->beta = 0.0;
      S->alpha = 0.0;
      S->pq = 0.0;
      S->iter = 0;
      S->status = 0;
      S->done = tot[NV - 1] == 0.0 ? 1 : 0;
      S->repl_norm = S->delta > 0.0 ? S->delta * sqrt(tot[NV - 1]) : 0.0;
      S->rho_slot[0] = tot[0];
      S->iter_base = 0;
      break;
    case kFinPcgAlpha: {
      const double pqt = tot[0];
      S->pq = pqt;
      const double a = S->rho / pqt;
      S->alpha = a;
      if (!(pqt > 0.0) || !isfinite(a)) {
        S->status = 2;
        S->done = 1;
      }
      break;
    }
    case kFinPcgBeta: {
      const double rho1 = tot[0];
      S->beta = rho1 / S->rho;
      S->rho = rho1;
      S->rr = tot[NV - 1];
      S->iter += 1;
      // the next batch starts here (fused batches read these)
      S->rho_slot[0] = rho1;
      S->iter_base = S->iter;
      const double rnorm = sqrt(tot[NV - 1]);
      if (!isfinite(rnorm) || !isfinite(rho1)) {
        S->status = 2;
        S->done = 1;
      } else if (rnorm < S->tol * sqrt(S->bnorm2) || rnorm == 0.0) {
        // (a zero residual is the exact solution, also for a zero right-hand side)
        S->status = 0;
        S->done = 1;
      } else if (S->iter > S->maxit) {
        S->status = 1;
        S->done = 1;
      } else if (rnorm < S->repl_norm) {
        S->status = 3;
        S->done = 1;
      }
      break;
    }
    case kFinReplace: {
      S->rho = tot[0];
      S->rr = tot[NV - 1];
      S->rho_slot[0] = tot[0];
      S->iter_base = S->iter;
      const double rnorm = sqrt(tot[NV - 1]);
      S->repl_norm = S->delta > 0.0 ? S->delta * rnorm : 0.0;
      if (!isfinite(rnorm)) {
        S->status = 2;
        S->done = 1;
      } else if (rnorm < S->tol * sqrt(S->bnorm2) || rnorm == 0.0) {
        S->status = 0;
        S->done = 1;
      } else {
        S->status = 0;
        S->done = 0;
      }
      break;
    }
    case kFinOc: {
      const double lam = 0.5 * (O->l1 + O->l0);
      O->vol = tot[0];
      if (tot[0] - O->g > 0.0) O->l0 = lam;
      else O->l1 = lam;
      O->counter += 1;
      O->lam_last = lam;
      const bool more = (O->l1 - O->l0) / (O->l1 + O->l0) > 1.0e-5 && O->maxiter > O->counter;
      O->done = more ? 0 : 1;
      break;
    }
    default:
      break;
  }
}

}  // namespace

void launch_finalize(int op, int nv, const double *partials, unsigned G, PcgScalars *S,
                     OcState *O, double *out, hipStream_t s) {
  if (nv == 1)
    hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(NT), 0, s, op, partials, G, S, O, out);
  else if (nv == 2)
    hipLaunchKernelGGL(k_finalize<2>, dim3(1), dim3(NT), 0, s, op, partials, G, S, O, out);
  else
    throw Error(kInternal, "finalize: unsupported width");
  HIPCHK(hipGetLastError());
}

}  // namespace icarat_hip
