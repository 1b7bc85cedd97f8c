// mixed.cpp -- mixed-precision solver (see mixed.hpp).
#include "mixed.hpp"

#include <cmath>

#include "context.hpp"
#include "mixed_kernels.hpp"

namespace icarat_hip {

MixedSolver::MixedSolver(Context &ctx) : ctx_(ctx), n3_(ctx.n3_) {
  ctx_.ensure_device();
  hipStream_t s = ctx_.stream();
  ke32_.alloc(576);
  {
    float k[576];
    const double *ke = ctx_.problem().ke0();
    for (int i = 0; i < 576; ++i) k[i] = (float)ke[i];
    HIPCHK(hipMemcpyAsync(ke32_.p, k, sizeof(k), hipMemcpyHostToDevice, s));
    ctx_.sync();
  }
  young32_.alloc(ctx_.lattice().nelx);
  for (DevBuf<float> *b : {&dinv32_, &r32_, &p32_[0], &p32_[1], &q32_, &d32_, &u32_, &y32_}) {
    b->alloc(n3_);
    HIPCHK(hipMemsetAsync(b->p, 0, b->bytes(), s));
  }
  ctx_.sync();
}

MixedSolver::~MixedSolver() {
  for (int m = 0; m < 2; ++m)
    for (int p = 0; p < 2; ++p) {
      if (exec_[m][p]) (void)hipGraphExecDestroy(exec_[m][p]);
      if (graph_[m][p]) (void)hipGraphDestroy(graph_[m][p]);
    }
}

void MixedSolver::prepare() {
  if (ready_) return;
  hipStream_t s = ctx_.stream();
  launch_f64_to_f32(ctx_.lattice().nelx, ctx_.d_young, young32_, s);
  launch_f64_to_f32(n3_, ctx_.d_dinv, dinv32_, s);
  ready_ = true;
}

void MixedSolver::apply(const double *u, double *y, int precision) {
  prepare();
  hipStream_t s = ctx_.stream();
  const int mode = precision == ICARAT_HIP_BF16 ? 1 : 0;
  launch_f64_to_f32(n3_, u, u32_, s);
  launch_apply_k_mixed(ctx_.lattice(), mode, ke32_, young32_, ctx_.d_free, u32_, y32_, s);
  launch_f32_to_f64(n3_, y32_, y, s);
}

void MixedSolver::build_graph(int mode) {
  if (exec_[mode][0]) return;
  hipStream_t s = ctx_.stream();
  const Lattice &L = ctx_.lattice();
  for (int par = 0; par < 2; ++par) {
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      for (int j = 0; j < kBatch; ++j) {
        float *p_old = p32_[(j + par) & 1], *p_new = p32_[(j + par + 1) & 1];
        launch_pcg_stencil_mixed(L, mode, ke32_, young32_, ctx_.d_free, r32_, dinv32_, p_old,
                                 p_new, q32_, ctx_.d_pcg, ctx_.d_partials, s);
        launch_pcg_update32(n3_, d32_, r32_, p_new, q32_, dinv32_, ctx_.d_pcg, ctx_.d_partials, s);
      }
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(s, &graph_[mode][par]));
    HIPCHK(hipGraphInstantiate(&exec_[mode][par], graph_[mode][par], nullptr, nullptr, 0));
  }
}

// Runs graph batches of the low-precision PCG until its stop flag rises.
void MixedSolver::iterate(int mode, int maxit) {
  build_graph(mode);
  hipStream_t s = ctx_.stream();
  PcgScalars *h = ctx_.h_pcg.p;
  const long long max_batches = (long long)(maxit - h->iter) / kBatch + 3;
  for (long long b = 0; b < max_batches && !h->done; ++b) {
    HIPCHK(hipGraphLaunch(exec_[mode][h->iter & 1], s));
    HIPCHK(hipMemcpyAsync(h, ctx_.d_pcg.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
    ctx_.sync();
  }
  if (!h->done) throw Error(kInternal, "low-precision PCG did not reach its stop test");
}

// Jacobi-PCG with reliable updates (residual replacement): the iteration runs
// in FP32 storage with the MFMA contraction; the solution accumulates in
// FP64, and whenever the recursive residual has dropped by `delta` since the
// last replacement (and at the stop test) the FP64 true residual b - K64 x
// replaces it while the search direction is kept.  The FP64 true residual is
// the stop criterion.
void MixedSolver::solve(const icarat_solve_opts &o, icarat_solve_stats &st) {
  if (o.precond != ICARAT_HIP_PRECOND_JACOBI)
    throw Error(kInvalidArgument, "unknown preconditioner");
  if (o.precision != ICARAT_HIP_FP32 && o.precision != ICARAT_HIP_BF16)
    throw Error(kInvalidArgument, "unknown precision");
  const int mode = o.precision == ICARAT_HIP_BF16 ? 1 : 0;
  const double delta = o.inner_tol > 0.0 && o.inner_tol < 1.0 ? o.inner_tol : 0.1;
  const int max_repl = o.max_refine > 0 ? o.max_refine : 200;
  prepare();
  hipStream_t s = ctx_.stream();
  const Lattice &L = ctx_.lattice();
  PcgScalars *h = ctx_.h_pcg.p;
  PcgScalars s0{};
  s0.tol = o.tol;
  s0.maxit = o.maxit;
  s0.delta = delta;
  *h = s0;
  HIPCHK(hipMemcpyAsync(ctx_.d_pcg.p, h, sizeof(PcgScalars), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemsetAsync(ctx_.d_x.p, 0, ctx_.d_x.bytes(), s));
  launch_replace32(n3_, ctx_.d_b, nullptr, r32_, d32_, p32_[0], dinv32_, ctx_.d_pcg,
                   ctx_.d_partials, 1, s);
  HIPCHK(hipMemcpyAsync(h, ctx_.d_pcg.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
  ctx_.sync();
  const double bnorm = std::sqrt(h->bnorm2);
  double best = bnorm;
  int repl = 0, no_gain = 0;
  bool fallback = false;
  st.status = ICARAT_HIP_OK;
  while (bnorm > 0.0) {
    iterate(mode, o.maxit);
    if (h->status == 2) {
      fallback = true;
      break;
    }
    if (h->status == 1) break;
    // reliable update: x64 += d32, r = b - K64 x64
    launch_ir_accumulate(n3_, d32_, ctx_.d_x, s);
    launch_apply_k64(L, ctx_.d_ke, ctx_.d_young, ctx_.d_free, ctx_.d_x, ctx_.d_y, s);
    launch_replace32(n3_, ctx_.d_b, ctx_.d_y, r32_, nullptr, nullptr, dinv32_, ctx_.d_pcg,
                     ctx_.d_partials, 0, s);
    HIPCHK(hipMemcpyAsync(h, ctx_.d_pcg.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
    ctx_.sync();
    ++repl;
    if (h->done) break;  // FP64 true residual below tol (or NaN: status 2)
    const double rn = std::sqrt(h->rr);
    if (rn < 0.5 * best) {
      best = rn;
      no_gain = 0;
    } else if (++no_gain >= 8) {
      fallback = true;  // the low-precision iteration has stalled
      break;
    }
    if (repl >= max_repl || h->iter > o.maxit) break;
  }
  int iters = h->iter;
  st.fallback = 0;
  if (bnorm > 0.0 && (fallback || h->status == 2) && iters < o.maxit) {
    // finish in FP64 from the current x64
    st.fallback = 1;
    launch_ir_accumulate(n3_, d32_, ctx_.d_x, s);
    const PcgScalars r = ctx_.pcg64(o.tol, o.maxit - iters, true);
    iters += r.iter;
    if (r.status == 2) throw Error(kNumerical, "PCG breakdown");
    *h = r;
  }
  st.iters = iters;
  st.refine_steps = repl;
  st.rec_relres = bnorm > 0.0 ? std::sqrt(h->rr) / bnorm : 0.0;
  if (bnorm > 0.0 && !(h->done && h->status == 0)) st.status = ICARAT_HIP_NOT_CONVERGED;
}

}  // namespace icarat_hip
