// multigrid.hpp -- geometric multigrid preconditioner for the voxel lattice:
// a hierarchy of 2x coarser voxel grids of the same box (the coarse Young's
// moduli are the means of their 8 children, the Dirichlet boxes are replayed
// so coarse constraints sit at the same coordinates), damped Jacobi smoothing
// with each level's own element stiffness, trilinear transfers, and a
// symmetric V-cycle as the preconditioner of (flexible) PCG with the FP64
// verified stop.
#pragma once
#include <memory>
#include <vector>

#include "../../include/icarat_hip.h"
#include "hip_util.hpp"

namespace icarat_hip {

class Context;

class Multigrid {
 public:
  Multigrid(Context &fine, const icarat_mg_opts &o);
  ~Multigrid();
  const icarat_mg_opts &opts() const { return o_; }
  int levels() const { return 1 + (int)lv_.size(); }
  Context &level(int l);
  // coarse Young's moduli and Jacobi diagonals after the fine density changed
  void refresh();
  // z = M(r) on the fine lattice (device pointers; z may not alias r)
  void vcycle(const double *r, double *z);
  // K u = b of the fine context (d_b -> d_x) by PCG with M = V-cycle
  void solve(const icarat_solve_opts &o, icarat_solve_stats &st);
  void drop_graphs();

 private:
  Context &f_;
  icarat_mg_opts o_;
  std::vector<std::unique_ptr<Context>> lv_;  // levels 1..L-1
  std::vector<double> omega_;                  // Jacobi damping per level
  DevBuf<double> z_, y0_, r0_, ones_;
  hipGraph_t graph_[2] = {};
  hipGraphExec_t exec_[2] = {};
  void cycle(int l, const double *b, double *x);
  void enqueue_batch(int par);
  void build_graph();
};

}  // namespace icarat_hip
