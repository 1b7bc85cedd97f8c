// multigrid.hpp -- geometric multigrid preconditioner for the voxel lattice:
// a hierarchy of 2x coarser voxel grids of the same box (the coarse Young's
// moduli are the means of their 8 children, the Dirichlet boxes are replayed
// so coarse constraints sit at the same coordinates), damped Jacobi smoothing
// with each level's own element stiffness, trilinear transfers, and a
// symmetric V-cycle as the preconditioner of (flexible) PCG with the FP64
// verified stop.
#pragma once
#include <algorithm>
#include <map>
#include <memory>
#include <vector>

#include "../../include/icarat_hip.h"
#include "hip_util.hpp"
#include "mg_kernels.hpp"

struct _rocblas_handle;  // rocBLAS handle type (rocblas/rocblas.h)

namespace icarat_hip {

class Context;

// the defaults of icarat_hip_mg_default_opts
inline icarat_mg_opts mg_default_opts() {
  icarat_mg_opts o{};
  o.max_levels = 10;
  o.nu_pre = 1;
  o.nu_post = 1;
  o.nu_coarse = 6;
  o.omega = 1.95;
  o.batch_iters = 4;
  o.dense_dofs = -1;
  o.direct_nodes = 262144;
  o.coarse_avg = 0;
  o.w_levels = -1;
  o.w_switch = 0.0;
  o.compensated = 1;
  return o;
}

class Multigrid {
 public:
  Multigrid(Context &fine, const icarat_mg_opts &o);
  ~Multigrid();
  const icarat_mg_opts &opts() const { return o_; }
  int levels() const { return 1 + (int)lv_.size(); }
  Context &level(int l);
  // coarse Young's moduli and Jacobi diagonals after the fine density changed
  void refresh();
  // K u = b of the fine context (d_b -> d_x) by PCG with M = V-cycle; warm:
  // from the guess in d_x instead of 0
  void solve(const icarat_solve_opts &o, icarat_solve_stats &st, bool warm = false);
  // The solve in pieces (the fine context's d_b must be ready), each ordered
  // after the work on stream s and before later work on s:
  // start: coarse operators refreshed if the densities changed, PCG state
  // initialised (from 0, or with warm from the guess in d_x: r = b - K x),
  // z = M(r) with beta = 0, scalars to the host copy once s has completed;
  // enqueue_batch(par): batch_iters iterations on the fine context's stream
  // (p buffers at parity par); verify: the FP64 true residual replaces the
  // recursive one and a restart z = M(r), beta = 0 follows, scalars to the
  // host copy once s has completed.
  void start(double tol, int maxit, hipStream_t s, bool warm = false);
  void enqueue_batch(int par);
  // enqueue_batch(par) through this hierarchy's captured graph (built on first
  // use) on the fine context's stream
  void launch_batch(int par);
  void build_graph();  // the two parity graphs (no-op if built or graphs are off)
  void verify(hipStream_t s);
  // After a verification's scalars reached the host copy: true when the solve
  // ends -- the verified residual passed the tolerance, or it stalled.  Near
  // the FP64 floor of a solution (|u| far above |f|, where rounding u itself
  // costs a fraction of the tolerance) the recursive and the verified
  // residual differ by a fixed amount: after a failed verification the next
  // phase stops on a tighter recursive target (the verified residual is still
  // judged against the tolerance).  The best verified iterate is kept, and
  // after kStallVerifications verifications without a new best (or
  // kMaxVerifications in all) it is put back in x and the host copy marked
  // capped (status 1, reported as not converged).
  bool after_verify();
  // throws if the last refresh's Cholesky factorization failed (call once the
  // fine context's stream has completed the factorization, e.g. after finish)
  void check_factor() const;
  // end of a solve: the new inverse in place (if still pending), the stream
  // synchronised, the factorization's status checked
  void finish();
  void drop_graphs();
  // FP32 fine level (solve precision ICARAT_HIP_FP32): the fine level's
  // smoothing sweeps, its residual and the transfers to and from level 1 run
  // on FP32 vectors with the FP32 K.u (young and KE0 in FP32, FP32
  // accumulation); the sweep that produces the V-cycle's result computes and
  // stores in FP64, and the coarse levels stay FP64.  Applies where the fine
  // level uses the z-marching K.u; needs nu_pre = nu_post = 1.  Captured
  // graphs are rebuilt when the mode changes.
  void set_fine_fp32(bool on);
  // the FP32 fine level runs in the first phase of a solve only
  bool fine_fp32() const { return fp32_ && f32cur_ && fine_fp32_capable(); }
  bool fine_fp32_on() const { return fp32_ && fine_fp32_capable(); }
  // whether the coarsest level is solved exactly (dense inverse)
  bool dense() const { return dense_; }
  // One V-cycle z = V(r) of the fine lattice vector r, enqueued on the fine
  // context's stream; returns the device buffer holding z (the V-cycle's
  // result buffer; the next V-cycle overwrites it).  The coarse operators must
  // be current (refresh_now).
  const double *vcycle(const double *r);
  // the buffer that holds the last V-cycle's result (inside multigrid PCG: the
  // V-cycle of the current residual at every batch boundary)
  const double *result() const { return zres_; }
  // coarse operators and the coarsest inverse of the current densities, in
  // place before the call returns to the stream's order (no overlapped swap)
  void refresh_now();
  int batch_iters() const { return std::max(1, o_.batch_iters); }
  int w_levels() const { return wlev_; }
  long long w_fallbacks() const { return w_fallbacks_; }
  // after a batch: a W-cycle phase that broke down (z.r <= 0) continues with
  // the V-cycle; true if the solve was resumed (scalars on the host again once
  // s_ext has synchronized)
  bool recover_w(hipStream_t s_ext);
  // whether level l smooths with the fused one-thread-per-node kernel
  bool direct(int l) const { return direct_[l] != 0; }
  // coarsest-level Cholesky factorizations so far, and those redone with the
  // diagonal shift after a breakdown
  long long factorizations() const { return n_fact_; }
  long long shifted() const { return n_shifted_; }

  // ---- the load cases of one problem in one solve ----
  // Whether solve_cases can run this hierarchy's cases together: FP64, the
  // z-marching K.u on the fine level and on every level not smoothed by the
  // direct kernel, no Galerkin operators, an even batch, no overlapped
  // factorization, and one cycle shape for the whole solve (the W-cycle, if
  // any, runs to the end: w_switch 0).
  bool cases_capable(const icarat_solve_opts &o) const;
  // K u_c = b_c for the nc right-hand sides stacked at B (case c at c x the
  // fine lattice's vector length) into X; guess[c] != 0: X holds case c's
  // starting guess (else it starts from 0).  Every case runs the arithmetic
  // of its own solve -- flexible PCG with the V- or W-cycle, the verified FP64
  // stop, the endgame target and the stall guard -- while each kernel of a
  // batch is launched once for the running cases (the operator's Young's
  // moduli and masks read once per launch), so each case's iterate equals
  // that of its own solve bit for bit.  A case whose W-cycle broke down (z.r
  // <= 0, which its own solve would resume with the V-cycle) is left to the
  // caller: redo[c] = 1, solve it on its own.
  void solve_cases(const icarat_solve_opts &o, int nc, const double *B, double *X,
                   const char *guess, icarat_solve_stats *st, std::vector<char> &redo);
  // solve_cases in pieces, for a caller that interleaves it with other work
  // (every launch on the fine context's stream): mc_begin sets the solve up
  // (returns with the stream synchronized); mc_poll, after a launch's scalars
  // reached the host (the stream synchronized), verifies the cases whose
  // batch ended and returns true once every case is finished; else
  // mc_build_graph captures the next batch (if graphs are on and it is new)
  // and mc_launch enqueues it with the scalars' copy to the host; mc_end ends
  // the solve (stats and redo as solve_cases).
  void mc_begin(const icarat_solve_opts &o, int nc, const double *B, double *X,
                const char *guess);
  bool mc_poll();
  bool mc_ready() const { return mc_.cnt > 0; }  // mc_poll has chosen the next batch
  void mc_build_graph();
  void mc_launch();
  void mc_end(icarat_solve_stats *st, std::vector<char> &redo);

 private:
  Context &f_;
  icarat_mg_opts o_;
  std::vector<std::unique_ptr<Context>> lv_;  // levels 1..L-1
  std::vector<double> omega_;                  // Jacobi damping per level
  DevBuf<double> z_, y0_, r0_;
  DevBuf<double> xbest_;  // best verified iterate of a solve that stalls at its floor
  DevBuf<double> xlo_;    // low part of the compensated iterate (unallocated: plain updates)
  static constexpr int kStallVerifications = 4, kMaxVerifications = 50;
  static constexpr double kF32Switch = 30.0;  // FP32 fine level off below this x tol
  double best_ = 0.0;
  int since_best_ = 0, verifications_ = 0;
  bool fp32_ = false;
  DevBuf<float> z32_;  // FP32 fine level: the iterate
  bool fine_fp32_capable() const;
  std::vector<char> direct_;
  // direct levels: the coarse correction and the first post-sweep in one
  // kernel (launch_mg_prolong_jacobi); ICARAT_HIP_MG_FUSE=0 keeps them apart
  bool fuse_ = true;
  // A new density's coarsest inverse is factored synchronously before the
  // solve.  ICARAT_HIP_MG_OVERLAP=1 factors it on the side stream while the
  // first batches run with the previous inverse: about 5% faster at 192x64x64,
  // but its results differ from run to run at rounding level (only while the
  // factorization runs concurrently with the solve's kernels; input snapshots
  // and guard zones around every device buffer ruled out input races and
  // out-of-bounds writes), so it is off by default.
  bool overlap_ = false;
  // W-cycle: coarse levels 1..wcur_ are cycled twice per visit; wcur_ is
  // o_.w_levels while the solve runs the W-cycle and 0 once it has switched
  int wcur_ = 0;
  int wlev_ = 0;        // the W-cycle depth of this hierarchy (o_.w_levels or automatic)
  long long w_fallbacks_ = 0;  // W-cycle phases that ended in a breakdown
  // The running solve is in its first phase: the W-cycle (wlev_ > 0) and the
  // FP32 fine level (fp32_), until the recursive residual falls below
  // w_switch x tol; then the FP64 V-cycle
  bool wmode_ = false;
  bool f32cur_ = false;  // the cycle being run or captured has the FP32 fine level
  bool two_phase() const { return wlev_ > 0 || (fp32_ && fine_fp32_capable()); }
  double tol_ = 0.0;    // the running solve's tolerance (the switch point)
  // Galerkin coarse operators (coarse_avg 5): explicit element matrices per
  // coarse level (kmat_[l], 576 per element), the per-element smoothing bounds
  // and their maximum per level; the damping is folded into the levels' dinv
  bool galerkin_ = false;
  std::vector<DevBuf<double>> kmat_;
  DevBuf<double> mu_, mumax_;
  // level l >= 1 operator with the fused consumer (MG_JACOBI / MG_RESIDUAL)
  void coarse_op(int l, int mode, const double *x, const double *b, double w, double *out,
                 hipStream_t s);
  bool dense_ = false;
  DevBuf<double> ainv_;  // dense inverse of the coarsest K
  DevBuf<double> dpart_;  // partial products of the dense solve's tiles
  // Factorization status per stream (index 0: the fine context's stream, 1:
  // the side stream), each with a pinned host copy, and one rocBLAS handle
  // (with its workspace) per stream, so the two factorizations share nothing.
  DevBuf<int> info_[2];
  int *h_info_[2] = {nullptr, nullptr};
  long long n_fact_ = 0, n_shifted_ = 0;
  // Overlapped refresh: a new density's coarsest inverse is factored into
  // anext_ on the side stream while the solve's first kSwapAfter batches run
  // with the previous inverse, then copied into ainv_ before the next batch.
  // The swap point is fixed, so every run takes the same path.
  static constexpr int kSwapAfter = 2;
  DevBuf<double> anext_;
  hipStream_t side_ = nullptr;
  hipEvent_t ev_side_ = nullptr, ev_fact_ = nullptr;
  bool have_inverse_ = false;  // ainv_ holds the inverse of this or an earlier density
  bool swap_pending_ = false;
  int batches_ = 0;            // batches launched since the solve's start
  // side = 0: on the fine context's stream with handle 0, else on side_
  void factor_into(double *dst, int side, double shift);
  void factor_checked(double *dst);
  void swap_inverse(hipStream_t s);  // ainv_ = anext_ once factored (if pending)
  void before_batch(hipStream_t s);  // the swap at its fixed batch
  ::_rocblas_handle *blas_[2] = {nullptr, nullptr};
  double *zres_ = nullptr;  // buffer that holds the fine V-cycle's result
  // captured batches [W-cycle][parity]
  hipGraph_t graph_[2][2] = {};
  hipGraphExec_t exec_[2][2] = {};
  hipGraphExec_t exec_for(int par) const { return exec_[wmode_ ? 1 : 0][par & 1]; }
  // Partials of z.r and z.q produced by the fine level's last sweep (count
  // stays 0 when that sweep cannot produce them).
  struct ZDot {
    const double *q;
    double *part;
    unsigned count;
  };
  // V-cycle on level l with right-hand side b; returns the buffer holding the
  // result.  The first sweep from zero is already in the level's x for l >= 1
  // (written by the restriction) and for l = 0 when x0_ready.  emit = false
  // only follows the buffers.
  double *cycle(int l, const double *b, bool emit, bool x0_ready = false, ZDot *zd = nullptr);
  void enqueue_restart();                      // restart flag, z = M(r), beta = 0
  // Load cases together: per level the stacked cycle vectors (case c at c x
  // the level's vector length; level 0: x, y, r only), the fine PCG vectors
  // and scalars per case, the partials of nc cases, and the captured batches
  // per range of running cases.
  struct Cases {
    int nc = 0;
    double *X = nullptr;  // the iterates the captured batches were built with
    const double *B = nullptr;
    std::vector<DevBuf<double>> b, x, y, r, u, nv;
    DevBuf<double> R, P[2], Q, xlo, xbest, yt;
    DevBuf<PcgScalars> S;
    PinnedBuf<PcgScalars> hS;
    DevBuf<double> pst, pup, pz, dpart;
    double *zres = nullptr;  // base of the fine cycle's result
    std::map<int, hipGraph_t> graph;  // key: first case x 256 + case count
    std::map<int, hipGraphExec_t> exec;
    // the running solve: per case 0 running, 1 finished, 2 left to the caller;
    // verifications, the stall guard's counters and best residual
    std::vector<int> state, repl, since, verifs;
    std::vector<double> best;
    std::vector<char> redo;
    long long batches = 0, max_batches = 0;
    int lo = 0, cnt = 0;  // the range of the next batch
    int wsave = 0;        // wcur_ outside the solve
  };
  Cases mc_;
  void mc_alloc(int nc);
  void mc_drop_graphs();
  // the cycle of cycle() for cases [c0, c0 + ncs); b and the result are the
  // bases of level l's stacked vectors
  double *cycle_mc(int l, const double *b, bool x0_ready, ZDot *zd, int c0, int ncs);
  void mc_enqueue_batch(int lo, int cnt);  // one batch of cases [lo, lo + cnt)
  void mc_restart(int lo, int cnt);        // restart flags, z = M(r), beta = 0
  // whether level l's V-cycle starts with the sweep x = omega D^-1 b (a
  // smoothed coarsest level, or any other level with pre-smoothing)
  bool starts_swept(int l) const;
  unsigned long long refreshed_ = ~0ULL;       // density version of the coarse operators
  hipEvent_t ev_[2] = {};                      // ordering between stream s and the fine stream
};

}  // namespace icarat_hip
