// context.hpp -- device state of one SIMP problem and the host orchestration
// of its kernels (solve, filter, OC, MMA, the design loop).
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/icarat_hip.h"
#include "hip_util.hpp"
#include "kernels.hpp"
#include "mma_kernels.hpp"
#include "problem.hpp"

namespace icarat_hip {

class DeviceGCMMA;
class DeviceMMA;
class HelmFilter;
class LearnedPrecond;
class Multigrid;

// addPlot of io/export_graph.hpp:56-70: the line "x y" appended to path, the
// file created afresh when is_first (config_io.cpp)
void add_plot(const char *path, double x, double y, bool is_first);

// the fixed device scratch of the no-allocation entry point (noalloc.hip):
// its per-device address is looked up by the device set-up of a context
void noalloc_init_device(int device);
long long noalloc_max_elements();

class Context {
 public:
  explicit Context(const GridSpec &g);
  ~Context();

  Problem &problem() { return prob_; }
  const Lattice &lattice() const { return L_; }
  hipStream_t stream() const { return stream_; }

  int make_dof();                          // numbering + device upload of masks
  // load cases: add_neumann adds to the current case; solve uses it; simp_run
  // sums compliance and sensitivities over all cases
  int new_load_case();
  void select_load_case(int i);
  void use_load_case(int i);               // d_f <- case i (device copy)
  void sync_fext();                        // host fext -> device lattice vector
  void mark_fext_dirty() { fext_dirty_ = true; solved_ = false; }
  void ensure_fext() { if (fext_dirty_) sync_fext(); }
  void set_density(const double *xphys);   // host -> device densities, young, dinv
  void set_density_device();               // d_dens already holds the densities

  // y = K u on numeq vectors (host pointers)
  void apply_k(const double *u, double *y, int precision);
  void get_diag(double *d);
  // FP64 / mixed solve of K u = fext; result kept on device (d_u, nodal with
  // Dirichlet values); optional numeq copy to host.
  void solve(const icarat_solve_opts &o, double *u_host, icarat_solve_stats *st);
  // Every load case solved together (FP64 Jacobi-PCG or multigrid PCG): one
  // multi-column iteration, each case with its own scalars and verified stop,
  // the same arithmetic as a solve of that case alone.  u_host: numeq values
  // per case (may be NULL); st[ncase] (may be NULL).  Afterwards d_u, d_x hold
  // the active case's solution.
  void solve_cases(const icarat_solve_opts &o, double *u_host, icarat_solve_stats *st);
  // o can be solved by one multi-column iteration (FP64 Jacobi or multigrid,
  // z-marching K.u; multigrid: Multigrid::cases_capable)
  bool cases_together(const icarat_solve_opts &o) const;
  // The design loop's automatic choice (case_mode 0) runs the cases together
  // with Jacobi-PCG up to this many elements (one launch over all cases beat
  // one launch per case on the batched workload's problems, 1.1-2x up to
  // 128x64x64, 0.9x at 96^3 and 160x80x80) and with multigrid PCG at every
  // size (1.3-2.0x from 32^3 to 160x80x80).
  static constexpr long long kCasesTogetherMaxElements = 600000;
  bool cases_auto(const icarat_solve_opts &o) const {
    return o.precond == ICARAT_HIP_PRECOND_MG || L_.nelx <= kCasesTogetherMaxElements;
  }
  // The multigrid solve of every case in pieces (Multigrid::mc_begin and
  // the following on this context's stream), for the batched engine:
  // cases_start (the stream synchronized on return); after each launch has
  // completed, cases_poll (true once every case is finished); else
  // cases_build_graph and cases_launch; then cases_finish (per-case stats, a
  // case left over by the W-cycle solved on its own, warm starts saved).
  void cases_start(const icarat_solve_opts &o);
  bool cases_poll();
  bool cases_ready() const;  // polled since the start, a batch to launch
  void cases_build_graph();
  void cases_launch();
  void cases_finish(const icarat_solve_opts &o, icarat_solve_stats *st);
  // after the cases' solve: each case's true residual and its compliance and
  // sensitivity terms, in case order
  void simp_cases_post(icarat_solve_stats *S);
  double true_residual_device();          // ||f - K d_u|| / ||f||
  double true_residual_host(const double *u);
  void compliance_sensitivity(double vnorm, double *c, double *dfds, double *dgds);
  // element-mean stress (nelx x 6, may be null) and von Mises stress of the last solve
  void mises_host(double *mises, double *stress);
  // as mises_host with the element stresses staged in the library's fixed
  // device scratch array instead of a per-call allocation (at most
  // noalloc_max_elements() elements when stress is requested)
  void mises_noalloc(double *mises, double *stress);
  // nodal values of the last solve (numnp x 3, node order; Dirichlet values
  // included, |v| < 1e-15 flushed as analysis/analysis_utility.hpp:26-28)
  void nodal_values_host(double *u);

  void filter_setup(double radius);
  void filter_device(const double *in, double *out, int is_sens);  // device pointers
  void filter_host(const double *in, double *out, int is_sens);
  // Filter::sensitivityFilter on host arrays (x, dfds -> out, nelx each)
  void sensitivity_filter_host(const double *x, const double *dfds, double *out);
  void threshold_host(const double *in, double beta, double T, double *out, int deriv);
  // Helmholtz PDE filter (FilterHelmholtz): set-up and both modes on host vectors
  void helm_setup(double radius);
  void helm_host(const double *in, double *out, int is_sens);
  int oc_device(double *x, const double *dc, const double *dv, double g, double xmin, double xmax,
                int use_projection, double beta);
  int oc_host(double *x, const double *dfdx, double g, const double *dgdx, double xmin,
              double xmax, int use_projection, double beta);

  void simp_run(const icarat_simp_opts &o, double *x, double *hist, icarat_simp_stats *st);
  // the loop in phases: begin; per step: pre (filter, density, volume), per load
  // case a solve then case_post (compliance, sensitivity), update (chain-rule
  // filters, OC / MMA); end
  void simp_begin(const icarat_simp_opts &o, const double *x);
  void simp_pre(int step);
  void simp_case_post(int lc, const icarat_solve_stats &ss);
  void simp_update(int step, double *hist);
  void simp_end(double *x, icarat_simp_stats *st);
  struct SimpState {
    icarat_simp_opts o{};
    icarat_simp_stats S{};
    double t_all = 0.0, V0 = 0.0, vol = 0.0, beta = 1.0, c = 0.0, relres = 0.0;
    double c_old = 0.0;  // the convergence test's previous compliance (Optimize::object_old)
    long long iters = 0;
  } simp_;

  // learned preconditioner: set-up, a harvesting Jacobi solve, z = M(r) on numeq vectors
  void learned_setup(const icarat_learned_opts &o);
  LearnedPrecond &learned();
  void learned_harvest(const icarat_solve_opts &o, double *u_host, icarat_solve_stats *st);
  void learned_apply(const double *r, double *z);
  void time_kernel(int kernel, int reps, double *avg_ms, double *flops, double *bytes);
  void set_stencil(int kernel, int sym_max);
  // kernel shape and z-chunk count of the FP64 K.u (keeps the symmetry class);
  // captured graphs are rebuilt
  void set_stencil_tuning(int kernel, int nchunks);
  void drop_graphs();

  // ---- device buffers (public for the solver components) ----
  Lattice L_;
  long long n3_ = 0;  // 3 * Np
  DevBuf<double> d_ke;                // KE0 (576)
  KeCoef kc_;                         // KE0 as the FP64 stencil kernels take it
  DevBuf<float> d_ke32;               // KE0 in FP32 for the MFMA contraction
  DevBuf<float> d_young32;            // SIMP young in FP32
  DevBuf<uint8_t> d_free;             // per node free mask
  DevBuf<int> d_dof;                  // numnp*3 reference DOF numbers
  DevBuf<double> d_dirval;            // numnp*3 Dirichlet values (AoS)
  DevBuf<double> d_dirlat;            // Dirichlet values on the lattice (SoA, 0 at free DOFs)
  DevBuf<double> d_f;                 // fext of the current load case on the lattice (0 at constrained DOFs)
  std::vector<DevBuf<double>> d_fcases_;  // every load case's fext on the lattice
  DevBuf<double> d_b;                 // right-hand side of the current solve
  DevBuf<double> d_dens, d_young;     // physical densities, SIMP young
  DevBuf<double> d_dinv;              // 1/diag K at free DOFs
  DevBuf<double> d_x, d_r, d_q, d_p[2];  // PCG vectors
  DevBuf<double> d_u;                 // nodal solution (free + Dirichlet values)
  DevBuf<double> d_nv;                // flushed nodal values for sensitivities
  DevBuf<double> d_y;                 // scratch lattice vector
  DevBuf<double> d_numeq[2];          // numeq staging
  DevBuf<double> d_partials;
  DevBuf<double> d_partials2;         // update partials of fused PCG batches
  DevBuf<unsigned> d_counter;
  DevBuf<PcgScalars> d_pcg;
  PinnedBuf<PcgScalars> h_pcg;
  DevBuf<double> d_scal;              // small reduction results
  PinnedBuf<double> h_scal;
  // state of the MMA / GCMMA dual iteration (device) and its host copy
  MmaDualState *mma_dual_state() { return (MmaDualState *)d_mmadual_.p; }
  MmaDualState *mma_dual_state_host() { return (MmaDualState *)h_mmadual_.p; }
  DevBuf<double> d_e[8];              // element work vectors
  DevBuf<OcState> d_oc;
  PinnedBuf<OcState> h_oc;
  DevBuf<int> d_flin;
  DevBuf<double> d_fw;
  FilterTable ftab_{0, 0, nullptr, nullptr};
  bool filter_ready_ = false;
  double radius_ = 0.0;
  bool density_ready_ = false;
  bool solved_ = false;
  bool fext_dirty_ = false;
  double vol_e_ = 1.0;
  DevBuf<double> d_mmadual_;
  PinnedBuf<double> h_mmadual_;
  DevBuf<double> d_sfilt_;  // 3 x nelx: x, dfds, out of the sensitivity filter
  std::string last_error;

  // PCG iteration batch captured as a graph, per contraction mode (FP64 VALU,
  // FP32 MFMA, BF16 MFMA) and parity of the p buffers
  static constexpr int kBatch = 16;
  hipGraphExec_t pcg_exec_[3][2] = {};
  hipGraph_t pcg_graph_[3][2] = {};
  // Graphs can be switched off (ICARAT_HIP_GRAPHS=0) for tools that cannot
  // trace graph launches; the kernels launched are the same.
  bool use_graphs_ = true;
  int replacements_ = 0;  // residual replacements of the last solve
  int fell_back_ = 0;     // the last solve switched to the FP64 contraction

  std::unique_ptr<DeviceMMA> mma_;
  std::unique_ptr<HelmFilter> helm_;
  std::unique_ptr<DeviceGCMMA> gcmma_;
  DeviceGCMMA &gcmma();                   // created on first use
  // GCMMA steps on host vectors: mode 0 OuterUpdate, 1 InnerUpdate (xmma in/out)
  void gcmma_host(int mode, double *xmma, double f0xnew, double fxnew, const double *xval,
                  double f0x, const double *df0dx, double fx, const double *dfdx, double xmin,
                  double xmax);
  // the GCMMA driver's loop (optimizer 2 of simp_run)
  void gcmma_run(const icarat_simp_opts &o, double *x, double *hist, icarat_simp_stats *st);
  // filter -> projection -> densities -> solve of every load case at the design
  // xd (device); returns the compliance summed over the cases and the volume
  // in *vol; with sens, also the sensitivities with the projection derivative
  // and the chain-rule filter applied (E_DFDS, E_DGDS)
  double gcmma_evaluate(const double *xd, bool sens, double *vol);
  // the design loop's filter (simp_.o.filter): cone or Helmholtz, device vectors
  void loop_filter(const double *in, double *out, int is_sens);
  // OC bisection whose volume callback filters with the Helmholtz filter
  int oc_device_helm(double *x, const double *dc, const double *dv, double g, double xmin,
                     double xmax, int use_projection, double beta);
  std::unique_ptr<LearnedPrecond> learned_;
  std::unique_ptr<Multigrid> mg_;
  void mg_setup(const icarat_mg_opts &o);
  // called by pcg64 after its set-up and after every batch of iterations
  std::function<void()> batch_hook_;

  // Jacobi-PCG on d_b -> d_x with FP64 vectors and scalars (from x = 0, or
  // from the current d_x when warm).  kmode selects the contraction of K.p:
  // -1 FP64 VALU, 0 FP32 MFMA, 1 BF16 MFMA.  delta > 0 replaces the recursive
  // residual by the FP64 true residual whenever it has dropped by delta; the
  // FP64 true residual is the stop test.  A low-precision contraction that
  // stalls is switched to FP64 in place.
  PcgScalars pcg64(double tol, int maxit, bool warm = false, int kmode = -1, double delta = 0.0);
  void assemble_nodal_solution();         // d_u = d_x + Dirichlet values
  void copy_out_solution(double *u_host); // numeq copy of d_u (if u_host), marks solved
  void prepare_rhs(hipStream_t s);        // d_b = f - K u_Dirichlet
  // FP64 PCG in pieces for a driver that interleaves several contexts: start
  // (rhs, init, scalars to h_pcg), verify (true residual replaces the
  // recursive one, scalars to h_pcg), finish (nodal solution, true residual,
  // stats).  The iterations themselves are enqueue_solver_batch /
  // launch_solver_batch.  The preconditioner is batch_precond_: Jacobi, or the
  // multigrid V-cycle (set up with the defaults on first use).
  void pcg_start(double tol, int maxit, hipStream_t s, bool warm = false);
  void pcg_verify(hipStream_t s);
  // once the verification's scalars are on the host: true if the solve ends
  // there (multigrid: passed or stalled at its floor; Jacobi: passed)
  bool pcg_after_verify();
  void pcg_finish(icarat_solve_stats &S);
  // a multigrid W-cycle phase that broke down resumes with the V-cycle
  // (enqueued on s, scalars to h_pcg); false: the breakdown stands
  bool pcg_recover(hipStream_t s);
  void enqueue_pcg_batch(int kmode, int par);  // kBatch Jacobi-PCG iterations on the stream
  // one batch of iterations of the interleaved solve (batch_precond_): enqueued
  // (for a caller's capture), or launched as this context's own graph
  void enqueue_solver_batch(int par);
  void launch_solver_batch(int par);
  void build_solver_graph();  // captures the graphs launch_solver_batch uses
  int batch_precond_ = ICARAT_HIP_PRECOND_JACOBI;
  // Warm starts: the last solution of each load case.  load_warm_guess copies
  // the active case's into d_x (false: none, d_x untouched); save_solution
  // records d_x for the active case.
  bool load_warm_guess(hipStream_t s);
  void save_solution(hipStream_t s);
  int active_case_ = 0;                   // load case whose fext is in d_f
  std::vector<DevBuf<double>> d_xcases_;
  std::vector<char> xcase_valid_;
  // bumped whenever the densities change (the multigrid hierarchy refreshes
  // its coarse operators once per version)
  unsigned long long density_version_ = 0;
  void sync() { HIPCHK(hipStreamSynchronize(stream_)); }
  // Device buffers are allocated on first device use, so mesh, numbering,
  // loads and KE0 work without a GPU.
  void ensure_device();

 private:
  Problem prob_;
  hipStream_t stream_ = nullptr;
  int device_ = 0;  // the device current at set-up (the stream's device)
  bool dev_ready_ = false;
  bool dof_uploaded_ = false;
  void upload_dof();
  void require_dof();
  void require_density();
  void build_pcg_graph(int kmode);
  void launch_pcg_batch(int kmode, int par);   // as a graph, or enqueued directly
  // load cases together: stacked vectors (case c at c * n3, p at c * 2 n3),
  // per-case scalars, partials and the captured batch
  struct Cases {
    int nc = 0;
    DevBuf<double> X, R, Q, P, B, pst, pup;
    DevBuf<PcgScalars> S;
    PinnedBuf<PcgScalars> hS;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  } mc_;
  void mc_alloc(int nc);
  void mc_enqueue_batch();
  void mc_drop_graph();
  // the device part of solve_cases; per-case stats without true residuals
  void mc_solve(const icarat_solve_opts &o, icarat_solve_stats *st);
  void mc_prepare();    // stacked buffers for the cases, right-hand sides in mc_.B
  void mc_save_warm();  // each case's solution as its warm start
  // d_x, d_u <- case c of the stacked solution, d_f <- its load
  void mc_take_case(int c);
  void simp_cases_together(const icarat_solve_opts &so);
};

}  // namespace icarat_hip
