// batch.cpp -- see batch.hpp.
#include "batch.hpp"

#include <algorithm>
#include <chrono>

#include "multigrid.hpp"

namespace icarat_hip {

namespace {
double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

Batch::Batch() { HIPCHK(hipStreamCreateWithFlags(&ms_, hipStreamNonBlocking)); }

Batch::~Batch() {
  if (ms_) (void)hipStreamSynchronize(ms_);
  for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
  if (fork_) (void)hipEventDestroy(fork_);
  if (ms_) (void)hipStreamDestroy(ms_);
}

void Batch::add(Context *c) {
  if (!c) throw Error(kInvalidArgument, "null context");
  if (std::find(ctx_.begin(), ctx_.end(), c) != ctx_.end())
    throw Error(kInvalidArgument, "context already in the batch");
  if (!c->problem().dof_ready()) throw Error(kNotReady, "make_dof has not been called");
  c->ensure_device();
  ctx_.push_back(c);
}

// One batch of iterations of every running problem: each problem's own
// captured graph on its own stream (separate hardware queues run them
// concurrently), after the work queued on the batch stream, with its scalars
// copied back; the batch stream then waits for every problem.
void Batch::launch_running(const std::vector<char> &run) {
  const int n = size();
  if (!fork_) HIPCHK(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
  while ((int)ev_.size() < n) {
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ev_.push_back(e);
  }
  // graphs are captured before any stream is made to wait
  for (int i = 0; i < n; ++i) {
    if (run[i] == kSingle) ctx_[i]->build_solver_graph();
    if (run[i] == kCases) ctx_[i]->cases_build_graph();
  }
  HIPCHK(hipEventRecord(fork_, ms_));
  for (int i = 0; i < n; ++i) {
    if (!run[i]) continue;
    Context &c = *ctx_[i];
    HIPCHK(hipStreamWaitEvent(c.stream(), fork_, 0));
    if (run[i] == kCases) {
      c.cases_launch();  // with the cases' scalars to the host
    } else {
      c.launch_solver_batch(c.h_pcg.p->iter & 1);
      HIPCHK(hipMemcpyAsync(c.h_pcg.p, c.d_pcg.p, sizeof(PcgScalars), hipMemcpyDeviceToHost,
                            c.stream()));
    }
    HIPCHK(hipEventRecord(ev_[i], c.stream()));
    HIPCHK(hipStreamWaitEvent(ms_, ev_[i], 0));
  }
}

void Batch::engine(double tol, int maxit, const std::function<int(int)> &start,
                   const std::function<void(int, const icarat_solve_stats &)> &done,
                   const std::function<void(int)> &done_cases) {
  const int n = size();
  std::vector<char> run(n, 0), admitted(n, 0), over(n, 0);
  // admission order: largest problem first
  std::vector<int> queue(n);
  for (int i = 0; i < n; ++i) queue[i] = i;
  std::stable_sort(queue.begin(), queue.end(),
                   [&](int a, int b) { return ctx_[a]->lattice().nelx > ctx_[b]->lattice().nelx; });
  size_t next = 0;
  long long load = 0;
  auto begin = [&](int i) {
    ctx_[i]->sync();
    run[i] = (char)start(i);
    if (run[i] == kCases && !done_cases) throw Error(kInternal, "batch engine: no cases handler");
    if (!run[i] && admitted[i] && !over[i]) {
      over[i] = 1;
      load -= ctx_[i]->lattice().nelx;
    }
  };
  auto admit = [&]() {
    while (next < queue.size()) {
      const int i = queue[next];
      const long long e = ctx_[i]->lattice().nelx;
      if (budget_ > 0 && load > 0 && load + e > budget_) break;
      ++next;
      admitted[i] = 1;
      load += e;
      begin(i);
    }
  };
  admit();
  HIPCHK(hipStreamSynchronize(ms_));
  for (;;) {
    // solves whose recursive residual passed: the FP64 true residual decides
    std::vector<int> ver, fin, rec;
    for (int i = 0; i < n; ++i) {
      if (!run[i]) continue;
      if (run[i] == kCases) {
        // the cases' own verifications (on the problem's stream)
        if (ctx_[i]->cases_poll()) fin.push_back(i);
        continue;
      }
      const PcgScalars &h = *ctx_[i]->h_pcg.p;
      if (h.done && h.status == 0) ver.push_back(i);
      else if (h.done && ctx_[i]->pcg_recover(ms_)) rec.push_back(i);  // W-cycle -> V-cycle
      else if (h.done) fin.push_back(i);  // iteration cap or breakdown
      else if (h.iter > maxit) {
        ctx_[i]->h_pcg.p->status = 1;
        fin.push_back(i);
      }
    }
    for (int i : ver) ctx_[i]->pcg_verify(ms_);
    if (!ver.empty() || !rec.empty()) HIPCHK(hipStreamSynchronize(ms_));
    for (int i : ver)
      if (ctx_[i]->pcg_after_verify()) fin.push_back(i);
    // finished solves: hand over, then the problem's next solve (if any)
    for (int i : fin) {
      if (run[i] == kCases) {
        run[i] = 0;
        done_cases(i);
      } else {
        icarat_solve_stats st{};
        ctx_[i]->pcg_finish(st);
        run[i] = 0;
        done(i, st);
      }
      begin(i);
    }
    admit();  // problems that finished their loop free budget
    if (!fin.empty() || next > 0) HIPCHK(hipStreamSynchronize(ms_));
    bool any = false;
    for (int i = 0; i < n; ++i) any = any || run[i];
    if (!any && next >= queue.size()) break;
    if (!any) continue;
    // a solve that has just started or verified may already be done (zero
    // right-hand side, or the true residual passed): handle it before launching
    bool pending = false;
    for (int i = 0; i < n; ++i) {
      if (run[i] == kSingle && ctx_[i]->h_pcg.p->done) pending = true;
      if (run[i] == kCases && !ctx_[i]->cases_ready()) pending = true;  // not polled yet
    }
    if (pending) continue;
    launch_running(run);
    ++launches_;
    HIPCHK(hipStreamSynchronize(ms_));
  }
}

void Batch::set_precond(const icarat_solve_opts &o) {
  if (o.precision != ICARAT_HIP_FP64 ||
      (o.precond != ICARAT_HIP_PRECOND_JACOBI && o.precond != ICARAT_HIP_PRECOND_MG))
    throw Error(kInvalidArgument, "batched solves are FP64 Jacobi-PCG or multigrid PCG");
  for (Context *c : ctx_) c->batch_precond_ = o.precond;
}

void Batch::solve(const icarat_solve_opts &o, icarat_solve_stats *st) {
  set_precond(o);
  if (!(o.tol > 0.0) || o.maxit < 0) throw Error(kInvalidArgument, "tol must be > 0 and maxit >= 0");
  const int n = size();
  std::vector<char> started(n, 0);
  std::vector<double> t0(n, 0.0);
  engine(
      o.tol, o.maxit,
      [&](int i) -> int {
        if (started[i]) return 0;
        started[i] = 1;
        t0[i] = now_s();
        ctx_[i]->pcg_start(o.tol, o.maxit, ms_, o.warm != 0);
        return kSingle;
      },
      [&](int i, const icarat_solve_stats &s) {
        ctx_[i]->copy_out_solution(nullptr);
        if (st) {
          st[i] = s;
          st[i].seconds = now_s() - t0[i];
        }
      });
}

void Batch::simp_run(const icarat_simp_opts &o, double *const *x, double *const *hist,
                     icarat_simp_stats *st) {
  if (o.optimizer == 2)
    throw Error(kInvalidArgument, "the batched loop runs OC or MMA (GCMMA: icarat_hip_simp_run)");
  set_precond(o.solve);
  if (!x) throw Error(kInvalidArgument, "null design array");
  const int n = size();
  for (int i = 0; i < n; ++i)
    if (!x[i]) throw Error(kInvalidArgument, "null design vector");
  for (int i = 0; i < n; ++i) ctx_[i]->simp_begin(o, x[i]);
  // per problem: design step and load case of the solve in flight
  std::vector<int> step(n, 1), lc(n, -1);
  std::vector<double> t_solve(n, 0.0);
  // multigrid problems with several load cases solve them together, as the
  // single-problem loop's automatic choice does (same results either way)
  auto together = [&](Context &c) {
    if (c.problem().num_load_cases() < 2 || o.solve.precond != ICARAT_HIP_PRECOND_MG ||
        o.case_mode == 1)
      return false;
    if (!c.mg_) c.mg_setup(mg_default_opts());
    return c.cases_together(o.solve);
  };
  engine(
      o.solve.tol, o.solve.maxit,
      [&](int i) -> int {
        Context &c = *ctx_[i];
        if (step[i] > o.n_iter || c.simp_.S.converged) return 0;
        const int nc = c.problem().num_load_cases();
        if (lc[i] < 0) {
          c.simp_pre(step[i]);  // a new design iteration
          if (together(c)) {
            lc[i] = nc - 1;
            t_solve[i] = now_s();
            c.cases_start(o.solve);
            return kCases;
          }
        }
        ++lc[i];
        if (nc > 1) c.use_load_case(lc[i]);
        t_solve[i] = now_s();
        c.pcg_start(o.solve.tol, o.solve.maxit, ms_, o.solve.warm != 0);
        return kSingle;
      },
      [&](int i, const icarat_solve_stats &s) {
        Context &c = *ctx_[i];
        c.simp_.S.solve_s += now_s() - t_solve[i];
        c.copy_out_solution(nullptr);
        c.simp_case_post(lc[i], s);
        const int nc = c.problem().num_load_cases();
        if (lc[i] + 1 < nc) return;  // the next case of this design iteration
        if (nc > 1) c.use_load_case(c.problem().load_case());
        c.simp_update(step[i], hist ? hist[i] : nullptr);
        ++step[i];
        lc[i] = -1;
      },
      [&](int i) {
        Context &c = *ctx_[i];
        std::vector<icarat_solve_stats> S(c.problem().num_load_cases());
        c.cases_finish(o.solve, S.data());
        c.simp_.S.solve_s += now_s() - t_solve[i];
        c.simp_cases_post(S.data());
        c.use_load_case(c.problem().load_case());
        c.simp_update(step[i], hist ? hist[i] : nullptr);
        ++step[i];
        lc[i] = -1;
      });
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
  // timed on Jacobi-PCG launches (the stencil configuration is shared)
  for (Context *c : ctx_) c->batch_precond_ = ICARAT_HIP_PRECOND_JACOBI;
  for (int cfg = 0; cfg < 4; ++cfg) {
    apply_config(cfg);
    // every problem iterates (unreachable tolerance, no cap in reach): the
    // launches time real iterations on the current densities
    for (Context *c : ctx_) c->sync();
    for (Context *c : ctx_) c->pcg_start(1e-300, 1 << 30, ms_);
    const std::vector<char> all(n, 1);
    launch_running(all);  // warm (graphs built)
    HIPCHK(hipStreamSynchronize(ms_));
    HIPCHK(hipEventRecord(a, ms_));
    for (int r = 0; r < reps; ++r) launch_running(all);
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
