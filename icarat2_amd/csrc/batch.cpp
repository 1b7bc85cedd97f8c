// batch.cpp -- see batch.hpp.
#include "batch.hpp"

#include <algorithm>
#include <chrono>

namespace icarat_hip {

namespace {
double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

Batch::Batch() { HIPCHK(hipStreamCreateWithFlags(&ms_, hipStreamNonBlocking)); }

Batch::~Batch() {
  if (ms_) (void)hipStreamSynchronize(ms_);
  drop_graph();
  for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
  if (ms_) (void)hipStreamDestroy(ms_);
}

void Batch::add(Context *c) {
  if (!c) throw Error(kInvalidArgument, "null context");
  if (std::find(ctx_.begin(), ctx_.end(), c) != ctx_.end())
    throw Error(kInvalidArgument, "context already in the batch");
  if (!c->problem().dof_ready()) throw Error(kNotReady, "make_dof has not been called");
  c->ensure_device();
  ctx_.push_back(c);
  drop_graph();
}

void Batch::drop_graph() {
  if (ge_) (void)hipGraphExecDestroy(ge_);
  if (g_) (void)hipGraphDestroy(g_);
  ge_ = nullptr;
  g_ = nullptr;
}

void Batch::build_graph() {
  if (ge_) return;
  const int n = size();
  while ((int)ev_.size() < 2 * n) {
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ev_.push_back(e);
  }
  HIPCHK(hipStreamBeginCapture(ms_, hipStreamCaptureModeGlobal));
  try {
    for (int i = 0; i < n; ++i) {
      // fork: problem i's iterations on its own stream after the batch stream
      HIPCHK(hipEventRecord(ev_[2 * i], ms_));
      HIPCHK(hipStreamWaitEvent(ctx_[i]->stream(), ev_[2 * i], 0));
      ctx_[i]->enqueue_pcg_batch(-1, 0);
      // join back into the batch stream
      HIPCHK(hipEventRecord(ev_[2 * i + 1], ctx_[i]->stream()));
      HIPCHK(hipStreamWaitEvent(ms_, ev_[2 * i + 1], 0));
    }
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(ms_, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HIPCHK(hipStreamEndCapture(ms_, &g_));
  HIPCHK(hipGraphInstantiate(&ge_, g_, nullptr, nullptr, 0));
}

void Batch::solve_set(const std::vector<int> &act, double tol, int maxit,
                      std::vector<icarat_solve_stats> &st) {
  const int n = size();
  std::vector<char> in(n, 0);
  for (int i : act) in[i] = 1;
  // every context finishes its own pending work before the batch stream uses it
  for (Context *c : ctx_) c->sync();
  for (int i = 0; i < n; ++i) {
    Context &c = *ctx_[i];
    if (in[i]) {
      c.pcg_start(tol, maxit, ms_);
    } else {
      // parked: its branch of the graph skips every kernel
      PcgScalars park{};
      park.done = 1;
      *c.h_pcg.p = park;
      HIPCHK(hipMemcpyAsync(c.d_pcg.p, c.h_pcg.p, sizeof(PcgScalars), hipMemcpyHostToDevice, ms_));
    }
  }
  HIPCHK(hipStreamSynchronize(ms_));
  build_graph();
  // 0 running, 1 finished
  std::vector<char> fin(n, 1);
  for (int i : act) fin[i] = 0;
  for (;;) {
    // solves whose recursive residual passed: the FP64 true residual decides
    std::vector<int> ver;
    for (int i : act) {
      if (fin[i]) continue;
      const PcgScalars &h = *ctx_[i]->h_pcg.p;
      if (h.done && h.status == 0) ver.push_back(i);
      else if (h.done) fin[i] = 1;  // iteration cap or breakdown
    }
    for (int i : ver) ctx_[i]->pcg_verify(ms_);
    if (!ver.empty()) HIPCHK(hipStreamSynchronize(ms_));
    for (int i : ver)
      if (ctx_[i]->h_pcg.p->done) fin[i] = 1;
    std::vector<int> run;
    for (int i : act) {
      if (fin[i]) continue;
      PcgScalars &h = *ctx_[i]->h_pcg.p;
      if (h.iter > maxit) {
        h.status = 1;
        fin[i] = 1;
        continue;
      }
      run.push_back(i);
    }
    if (run.empty()) break;
    // the graph starts every branch at p-buffer parity 0
    for (int i : run) {
      Context &c = *ctx_[i];
      if (c.h_pcg.p->iter & 1)
        HIPCHK(hipMemcpyAsync(c.d_p[0].p, c.d_p[1].p, c.d_p[0].bytes(), hipMemcpyDeviceToDevice, ms_));
    }
    HIPCHK(hipGraphLaunch(ge_, ms_));
    ++launches_;
    for (int i : run) {
      Context &c = *ctx_[i];
      HIPCHK(hipMemcpyAsync(c.h_pcg.p, c.d_pcg.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, ms_));
    }
    HIPCHK(hipStreamSynchronize(ms_));
  }
  st.assign(n, icarat_solve_stats{});
  for (int i : act) ctx_[i]->pcg_finish(st[i]);
}

void Batch::solve(const icarat_solve_opts &o, icarat_solve_stats *st) {
  if (o.precision != ICARAT_HIP_FP64 || o.precond != ICARAT_HIP_PRECOND_JACOBI)
    throw Error(kInvalidArgument, "batched solves are FP64 Jacobi-PCG");
  if (!(o.tol > 0.0) || o.maxit < 0) throw Error(kInvalidArgument, "tol must be > 0 and maxit >= 0");
  std::vector<int> act(size());
  for (int i = 0; i < size(); ++i) act[i] = i;
  const double t0 = now_s();
  std::vector<icarat_solve_stats> s;
  solve_set(act, o.tol, o.maxit, s);
  const double dt = now_s() - t0;
  for (int i = 0; i < size(); ++i) {
    s[i].seconds = dt;
    ctx_[i]->copy_out_solution(nullptr);
    if (st) st[i] = s[i];
  }
}

void Batch::simp_run(const icarat_simp_opts &o, double *const *x, double *const *hist,
                     icarat_simp_stats *st) {
  if (o.solve.precision != ICARAT_HIP_FP64 || o.solve.precond != ICARAT_HIP_PRECOND_JACOBI)
    throw Error(kInvalidArgument, "batched solves are FP64 Jacobi-PCG");
  if (!x) throw Error(kInvalidArgument, "null design array");
  const int n = size();
  for (int i = 0; i < n; ++i) {
    if (!x[i]) throw Error(kInvalidArgument, "null design vector");
    ctx_[i]->simp_begin(o, x[i]);
  }
  int maxcase = 0;
  for (Context *c : ctx_) maxcase = std::max(maxcase, c->problem().num_load_cases());
  std::vector<icarat_solve_stats> s;
  for (int step = 1; step <= o.n_iter; ++step) {
    for (Context *c : ctx_) c->simp_pre(step);
    for (int lc = 0; lc < maxcase; ++lc) {
      std::vector<int> act;
      for (int i = 0; i < n; ++i) {
        const int nc = ctx_[i]->problem().num_load_cases();
        if (nc <= lc) continue;
        if (nc > 1) ctx_[i]->use_load_case(lc);
        act.push_back(i);
      }
      const double t0 = now_s();
      solve_set(act, o.solve.tol, o.solve.maxit, s);
      const double dt = now_s() - t0;
      for (int i : act) {
        ctx_[i]->simp_.S.solve_s += dt;
        ctx_[i]->copy_out_solution(nullptr);
        ctx_[i]->simp_case_post(lc, s[i]);
      }
    }
    for (int i = 0; i < n; ++i) {
      Context &c = *ctx_[i];
      if (c.problem().num_load_cases() > 1) c.use_load_case(c.problem().load_case());
      c.simp_update(step, hist ? hist[i] : nullptr);
    }
  }
  for (int i = 0; i < n; ++i) ctx_[i]->simp_end(x[i], st ? st + i : nullptr);
}

// configuration cfg: 0..2 z-marching with the default z-chunk count divided by
// 1, 2, 4; 3 the 3-D tile kernel
void Batch::apply_config(int cfg) {
  for (Context *c : ctx_) {
    if (cfg == 3)
      c->set_stencil_tuning(1, 0);
    else
      c->set_stencil_tuning(0, std::max(1, stencil_nchunks(c->lattice()) >> cfg));
  }
  drop_graph();
}

int Batch::autotune(int reps, double *ms_out) {
  if (reps <= 0) throw Error(kInvalidArgument, "reps must be positive");
  const int n = size();
  if (n == 0) throw Error(kNotReady, "empty batch");
  double best = 0.0;
  int best_cfg = 0;
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  for (int cfg = 0; cfg < 4; ++cfg) {
    apply_config(cfg);
    // every problem iterates (unreachable tolerance, no cap in reach): the
    // launches time real iterations on the current densities
    for (Context *c : ctx_) c->sync();
    for (Context *c : ctx_) c->pcg_start(1e-300, 1 << 30, ms_);
    build_graph();
    HIPCHK(hipGraphLaunch(ge_, ms_));  // warm
    HIPCHK(hipEventRecord(a, ms_));
    for (int r = 0; r < reps; ++r) HIPCHK(hipGraphLaunch(ge_, ms_));
    HIPCHK(hipEventRecord(b, ms_));
    HIPCHK(hipEventSynchronize(b));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, a, b));
    const double per = (double)t / reps;
    if (ms_out) ms_out[cfg] = per;
    if (cfg == 0 || per < best) {
      best = per;
      best_cfg = cfg;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  apply_config(best_cfg);
  for (Context *c : ctx_) c->solved_ = false;
  return best_cfg;
}

}  // namespace icarat_hip
