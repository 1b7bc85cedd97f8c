// learned.hpp -- the learned preconditioner: a 3-D U-Net on a 2x coarsened
// cell grid maps (stiffness feature, base-preconditioned residual) to the part
// of the error that the base preconditioner misses; flexible PCG uses
//     z = B(r) + sigma * P(net(feature, R(B(r)) / sigma)),
// B = D^-1 (Jacobi base, Jacobi-PCG fallback) or B = one multigrid V-cycle
// (multigrid base, multigrid-PCG fallback).  The network is trained online
// (Adam, on the project's own forward/backward kernels) on pairs harvested from
// the base method's PCG iterates of the first design iterations: input
// R(B(r_k)), target R(x* - x_k - B(r_k)) (both over sigma_k = RMS R(B(r_k))).
#pragma once
#include <memory>
#include <vector>

#include "../../include/icarat_hip.h"
#include "hip_util.hpp"
#include "unet.hpp"

namespace icarat_hip {

class Context;

class LearnedPrecond {
 public:
  LearnedPrecond(Context &ctx, const icarat_learned_opts &o);
  ~LearnedPrecond();
  const icarat_learned_opts &opts() const { return o_; }
  UNet &net() { return *net_; }
  void drop_graphs();  // captured batches (rebuilt at the next solve)
  // explicit parameters: the network is used as given (gain 1 until a harvest sets it)
  void set_params(const float *host);
  const NetGrid &grid() const { return g0_; }

  // Training data: snapshots of (x_k, r_k) taken every few batches of a
  // Jacobi solve; after the solve, targets from the converged x.
  void begin_harvest();
  void snapshot();            // at a batch boundary of the running solve
  void end_harvest(bool converged);  // targets need the converged x
  int samples() const;
  int harvested() const { return harvested_; }  // harvesting solves finished
  // Adam steps on random harvested samples; returns the mean relative squared
  // error ||net - target||^2 / ||target||^2 of the last (up to 20) steps, -1
  // without samples.
  double train(int steps);
  double last_loss() const { return last_loss_; }
  double val_loss() const { return val_loss_; }
  double gain() const { return gain_; }
  int fallback_iter() const { return fallback_iter_; }  // iteration of the last fallback
  // The design loop's usefulness gate: the first solve with the network is a
  // trial against the iterations of the last harvesting (base-method) solve;
  // a network that needs more is switched off for the rest of the loop.
  void note_base_iters(int iters) { base_iters_ = iters; }
  bool trial_pending() const { return active_ && !trialed_ && base_iters_ > 0; }
  void trial_result(int iters) {
    trialed_ = true;
    trial_iters_ = iters;
    if (iters > base_iters_) active_ = false;
  }
  int trial_iters() const { return trial_iters_; }
  int trial_base_iters() const { return trialed_ ? base_iters_ : -1; }
  bool active() const { return active_; }
  long long total_steps() const { return total_steps_; }
  // Flexible PCG with the network preconditioner on ctx.d_b -> ctx.d_x.
  void solve(const icarat_solve_opts &o, icarat_solve_stats &st);
  // z = M(r) once (for tests): r, z lattice vectors on the device
  void apply(const double *r, double *z);

 private:
  Context &ctx_;
  icarat_learned_opts o_;
  bool mg_base_ = false;           // base preconditioner: the multigrid V-cycle
  // B(r) on the device: D^-1 r is applied inside the restriction and the
  // prolongation (dinv operand); the V-cycle result is passed with ones_
  const double *base_vec(const double *r, const double **dinv);
  NetGrid g0_;
  std::unique_ptr<UNet> net_;
  long long ncell_;
  DevBuf<float> feat_, s3_, out_;
  DevBuf<__bf16> xin_;
  DevBuf<double> scale_, z_, ones_;
  DevBuf<double> corr_;  // the network's correction alone (descent guard)
  // sample ring: valid_[k] once slot k holds a finished (input, target) pair
  int cap_ = 0, next_ = 0;
  std::vector<char> valid_;
  DevBuf<__bf16> sx_;            // cap x ncell x 8
  DevBuf<float> szr_, st_;       // cap x ncell x 3 (scaled residual part; R x_k then target)
  DevBuf<double> ssig_;          // cap sigmas
  std::vector<int> pending_;     // slots filled by the running harvest
  DevBuf<float> rxs_;
  bool harvesting_ = false;
  int harvest_batches_ = 0;
  int harvested_ = 0;
  double last_loss_ = -1.0;
  double val_loss_ = -1.0;
  double gain_ = 1.0;     // targets are stored divided by gain_; the correction is scaled by it
  bool gain_set_ = false; // gain_ comes from the first harvested solve
  bool active_ = false;   // the network passed the gate
  bool graph_active_ = false;
  double graph_gain_ = 0.0;
  long long total_steps_ = 0;
  int fallback_iter_ = -1;
  int base_iters_ = 0, trial_iters_ = -1;
  bool trialed_ = false;
  uint64_t rng_;
  // graphs of batches of flexible-PCG iterations [kmode + 1][parity]
  hipGraph_t graph_[3][2] = {};
  hipGraphExec_t exec_[3][2] = {};
  void build_graph(int kmode);
  double evaluate();  // mean relative squared error over the stored samples
  void enqueue_batch(int kmode, int par);
  void precondition(const double *r);  // z_ = M(r)
};

}  // namespace icarat_hip
