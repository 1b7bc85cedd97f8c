// batch.hpp -- several independent SIMP problems (contexts of any size) driven
// together.  The PCG iterations of all of them are captured into one graph:
// one branch per problem on that problem's stream, forked from and joined into
// the batch stream through events, so the problems' kernels run concurrently
// and small problems share the GPU.  The graph is launched until every
// problem's solve has passed its FP64 verified stop; problems that are done
// (or not taking part in a round) skip their kernels.
#pragma once
#include <hip/hip_runtime.h>

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

  // one FP64 Jacobi-PCG solve of every context's current load case
  void solve(const icarat_solve_opts &o, icarat_solve_stats *st);
  // the SIMP loop of every context in lock step; per design iteration, load
  // case k of every problem that has one is solved in round k
  void simp_run(const icarat_simp_opts &o, double *const *x, double *const *hist,
                icarat_simp_stats *st);
  // Times whole-batch graph launches under each stencil configuration (applied
  // to every context: z-marching with 1/1, 1/2 or 1/4 of the default z-chunks,
  // or 3-D tiles) and keeps the fastest.  Needs densities on every context.
  // Returns the chosen configuration and the time of one launch per
  // configuration (ms[4]).
  int autotune(int reps, double *ms);

 private:
  std::vector<Context *> ctx_;
  hipStream_t ms_ = nullptr;
  std::vector<hipEvent_t> ev_;
  hipGraph_t g_ = nullptr;
  hipGraphExec_t ge_ = nullptr;
  long long launches_ = 0;
  void drop_graph();
  void build_graph();
  // Solves the contexts listed in act concurrently; the others are parked.
  void solve_set(const std::vector<int> &act, double tol, int maxit,
                 std::vector<icarat_solve_stats> &st);
  void apply_config(int cfg);
};

}  // namespace icarat_hip
