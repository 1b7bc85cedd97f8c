// batch.hpp -- several independent SIMP problems (contexts of any size) driven
// together.  Each problem advances through its own design loop; whenever
// problems have a solve running, their PCG iterations are launched together:
// each running problem's own captured graph on that problem's stream, forked
// from and joined into the batch stream through events, so the problems'
// kernels run concurrently on separate hardware queues and small problems
// share the GPU with large ones.  Every solve keeps its own FP64 verified stop.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <string>
#include <vector>

#include "../../include/icarat_hip.h"
#include "context.hpp"

namespace icarat_hip {

class Batch {
 public:
  Batch();
  ~Batch();
  Batch(const Batch &) = delete;
  Batch &operator=(const Batch &) = delete;

  void add(Context *c);
  int size() const { return (int)ctx_.size(); }
  long long graph_launches() const { return launches_; }

  // one FP64 solve of every context's current load case (Jacobi-PCG or
  // multigrid PCG, o.precond)
  void solve(const icarat_solve_opts &o, icarat_solve_stats *st);
  // the SIMP loop of every context; each problem advances on its own (design
  // phases between its solves), the solves of all problems share the launches
  void simp_run(const icarat_simp_opts &o, double *const *x, double *const *hist,
                icarat_simp_stats *st);
  // Times whole-batch launches under each stencil configuration (applied
  // to every context: z-marching with 1/1, 1/2 or 1/4 of the default z-chunks,
  // or 3-D tiles) and keeps the fastest.  Needs densities on every context.
  // Returns the chosen configuration and the time of one launch per
  // configuration (ms[4]).
  int autotune(int reps, double *ms);
  // Admission: problems enter the running set largest first while the running
  // problems' elements stay within the budget (at least one runs; 0: no
  // limit); an admitted problem runs its whole loop.
  void set_budget(long long elements) { budget_ = elements; }
  long long budget() const { return budget_; }

 private:
  std::vector<Context *> ctx_;
  hipStream_t ms_ = nullptr;
  std::vector<hipEvent_t> ev_;  // per problem: its batch of iterations done
  hipEvent_t fork_ = nullptr;   // the batch stream's work before the launches
  long long launches_ = 0;
  // no limit: on the 24-problem workload all problems at once ran 1.3x faster
  // than one after another, against 1.2x with a 3M-element budget and 1.1x
  // with 1.5M
  long long budget_ = 0;
  // run[i]: 0 idle, kSingle a single solve, kCases the multigrid solve of
  // all of problem i's load cases
  static constexpr char kSingle = 1, kCases = 2;
  void launch_running(const std::vector<char> &run);
  // The solve engine: start(i) begins problem i's next solve and returns its
  // kind -- kSingle (pcg_start on the batch stream), kCases (cases_start), or
  // 0 when problem i has no more solves; done(i, st) receives a finished
  // single solve, done_cases(i) a finished cases' solve (cases_finish is
  // left to it).  Iterates every running solve in shared launches until no
  // problem starts another solve.
  void engine(double tol, int maxit, const std::function<int(int)> &start,
              const std::function<void(int, const icarat_solve_stats &)> &done,
              const std::function<void(int)> &done_cases = {});
  void apply_config(int cfg);
  void set_precond(const icarat_solve_opts &o);
};

}  // namespace icarat_hip
