// context.cpp -- device state and host orchestration of one SIMP problem.
#include "context.hpp"
#include "learned_kernels.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>

#include "learned.hpp"
#include "multigrid.hpp"
#include "mixed_kernels.hpp"
#include "gcmma.hpp"
#include "helm.hpp"
#include "mma.hpp"

namespace icarat_hip {

namespace {
double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}
template <class T>
void h2d(T *dst, const T *src, size_t n, hipStream_t s) {
  if (n) HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
}
template <class T>
void d2h(T *dst, const T *src, size_t n, hipStream_t s) {
  if (n) HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
}
template <class T>
void d2d(T *dst, const T *src, size_t n, hipStream_t s) {
  if (n) HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, s));
}
}  // namespace

Context::Context(const GridSpec &g) : prob_(g) {
  L_ = prob_.lattice();
  n3_ = 3 * L_.Np;
  vol_e_ = prob_.element_volume();
}

void Context::ensure_device() {
  if (dev_ready_) {
    if (prob_.dof_ready() && !dof_uploaded_) upload_dof();
    return;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    throw Error(kHipError, "no HIP device available");
  HIPCHK(hipGetDevice(&device_));
  noalloc_init_device(device_);
  HIPCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  d_ke.alloc(576);
  h2d(d_ke.p, prob_.ke0(), 576, stream_);
  kc_ = make_ke_coef(prob_.ke0(), d_ke.p);
  if (const char *g = std::getenv("ICARAT_HIP_GRAPHS")) use_graphs_ = std::string(g) != "0";
  {
    float k32[576];
    for (int i = 0; i < 576; ++i) k32[i] = (float)prob_.ke0()[i];
    d_ke32.alloc(576);
    h2d(d_ke32.p, k32, 576, stream_);
    sync();
  }
  kc_.ke32 = d_ke32.p;
  for (DevBuf<double> *b : {&d_f, &d_b, &d_dinv, &d_x, &d_r, &d_q, &d_p[0], &d_p[1], &d_u, &d_nv,
                            &d_y, &d_dirlat}) {
    b->alloc(n3_);
    HIPCHK(hipMemsetAsync(b->p, 0, b->bytes(), stream_));
  }
  d_dens.alloc(L_.nelx);
  d_young.alloc(L_.nelx);
  d_young32.alloc(L_.nelx);
  for (auto &b : d_e) {
    b.alloc(L_.nelx);
    HIPCHK(hipMemsetAsync(b.p, 0, b.bytes(), stream_));
  }
  const long long filter_blocks = (long long)((L_.nx + 31) / 32) * ((L_.ny + 7) / 8) * ((L_.nz + 3) / 4);
  const long long nparts = std::max({stencil_blocks(L_), filter_blocks, 4096LL}) * 4;
  d_partials.alloc(nparts);
  d_partials2.alloc(2 * 4096);
  d_counter.alloc(4);
  HIPCHK(hipMemsetAsync(d_counter.p, 0, d_counter.bytes(), stream_));
  d_pcg.alloc(1);
  h_pcg.alloc(1);
  d_scal.alloc(16);
  h_scal.alloc(16);
  static_assert(sizeof(MmaDualState) % sizeof(double) == 0, "MmaDualState size");
  d_mmadual_.alloc(sizeof(MmaDualState) / sizeof(double));
  h_mmadual_.alloc(sizeof(MmaDualState) / sizeof(double));
  d_oc.alloc(1);
  h_oc.alloc(1);
  sync();
  dev_ready_ = true;
  if (prob_.dof_ready()) upload_dof();
}

Context::~Context() {
  if (!dev_ready_) return;
  for (int m = 0; m < 3; ++m)
    for (int k = 0; k < 2; ++k) {
      if (pcg_exec_[m][k]) (void)hipGraphExecDestroy(pcg_exec_[m][k]);
      if (pcg_graph_[m][k]) (void)hipGraphDestroy(pcg_graph_[m][k]);
    }
  mc_drop_graph();
  learned_.reset();
  mg_.reset();
  mma_.reset();
  if (stream_) (void)hipStreamSynchronize(stream_);
  // buffers release themselves; the stream goes last
  d_ke.release();
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Context::require_dof() {
  if (!prob_.dof_ready()) throw Error(kNotReady, "make_dof has not been called");
  ensure_device();
}
void Context::require_density() {
  require_dof();
  if (!density_ready_) throw Error(kNotReady, "set_density has not been called");
}

int Context::make_dof() {
  const int numeq = prob_.make_dof();
  dof_uploaded_ = false;
  density_ready_ = false;
  solved_ = false;
  fext_dirty_ = true;
  if (dev_ready_) upload_dof();
  return numeq;
}

void Context::upload_dof() {
  const int numeq = prob_.numeq();
  d_free.alloc(L_.numnp);
  h2d(d_free.p, prob_.free_mask().data(), L_.numnp, stream_);
  {
    // a node column x = 0 without free DOFs (a clamped face) is left to the
    // z-marching K.u's first tile column, whose tiles then start at x = 1
    const auto &fm = prob_.free_mask();
    bool none = L_.NX > 1;
    for (long long k = 0; k < (long long)L_.NY * L_.NZ && none; ++k) none = fm[k * L_.NX] == 0;
    kc_.xlo = none ? 1 : 0;
    drop_graphs();
  }
  d_dof.alloc(L_.numnp * 3);
  h2d(d_dof.p, prob_.dof().data(), L_.numnp * 3, stream_);
  d_dirval.alloc(L_.numnp * 3);
  h2d(d_dirval.p, prob_.dirichlet_values().data(), L_.numnp * 3, stream_);
  // Dirichlet values on the lattice: 0 at free DOFs
  {
    std::vector<double> lat(n3_, 0.0);
    const auto &dof = prob_.dof();
    const auto &dv = prob_.dirichlet_values();
    for (long long n = 0; n < L_.numnp; ++n)
      for (int c = 0; c < 3; ++c)
        if (dof[3 * n + c] >= numeq) lat[c * L_.Np + n] = dv[3 * n + c];
    h2d(d_dirlat.p, lat.data(), n3_, stream_);
    sync();
  }
  d_numeq[0].alloc(std::max(numeq, 1));
  d_numeq[1].alloc(std::max(numeq, 1));
  dof_uploaded_ = true;
  sync_fext();
}

void Context::sync_fext() {
  require_dof();
  const int numeq = prob_.numeq();
  const int nc = prob_.num_load_cases();
  if ((int)d_fcases_.size() != nc) {
    d_fcases_.clear();
    d_fcases_.resize(nc);
    for (auto &b : d_fcases_) {
      b.alloc(n3_);
      // the conversion below writes the real nodes only: the padding stays 0
      HIPCHK(hipMemsetAsync(b.p, 0, b.bytes(), stream_));
    }
  }
  // constrained entries of the lattice loads are 0 by construction
  for (int i = 0; i < nc; ++i) {
    h2d(d_numeq[0].p, prob_.case_fext(i).data(), numeq, stream_);
    launch_numeq_to_lattice(L_, d_dof, numeq, d_numeq[0], nullptr, d_fcases_[i], stream_);
  }
  d2d(d_f.p, d_fcases_[prob_.load_case()].p, n3_, stream_);
  active_case_ = prob_.load_case();
  sync();
  solved_ = false;
  fext_dirty_ = false;
}

void Context::use_load_case(int i) {
  ensure_fext();
  if (i < 0 || i >= (int)d_fcases_.size()) throw Error(kInvalidArgument, "no such load case");
  d2d(d_f.p, d_fcases_[i].p, n3_, stream_);
  active_case_ = i;
  solved_ = false;
}

bool Context::load_warm_guess(hipStream_t s) {
  const int c = active_case_;
  if (c >= (int)xcase_valid_.size() || !xcase_valid_[c]) return false;
  d2d(d_x.p, d_xcases_[c].p, n3_, s);
  return true;
}

void Context::save_solution(hipStream_t s) {
  const int c = active_case_;
  if (c >= (int)d_xcases_.size()) {
    d_xcases_.resize(c + 1);
    xcase_valid_.resize(c + 1, 0);
  }
  if (!d_xcases_[c].p) d_xcases_[c].alloc(n3_);
  d2d(d_xcases_[c].p, d_x.p, n3_, s);
  xcase_valid_[c] = 1;
}

int Context::new_load_case() {
  const int i = prob_.new_load_case();
  mark_fext_dirty();
  return i;
}

void Context::select_load_case(int i) {
  prob_.select_load_case(i);
  if (fext_dirty_ || !dev_ready_) {
    solved_ = false;
    return;
  }
  use_load_case(i);
}

void Context::set_density(const double *xphys) {
  require_dof();
  h2d(d_dens.p, xphys, L_.nelx, stream_);
  set_density_device();
  sync();
}

void Context::set_density_device() {
  require_dof();
  const auto &g = prob_.spec();
  launch_young(d_dens, d_young, L_.nelx, g.young1, g.young2, g.penalty, stream_);
  launch_f64_to_f32(L_.nelx, d_young, d_young32, stream_);
  launch_diag_inv(L_, d_ke, d_young, d_free, d_dinv, nullptr, stream_);
  density_ready_ = true;
  solved_ = false;
  ++density_version_;
}

void Context::apply_k(const double *u, double *y, int precision) {
  require_density();
  const int numeq = prob_.numeq();
  h2d(d_numeq[0].p, u, numeq, stream_);
  launch_numeq_to_lattice(L_, d_dof, numeq, d_numeq[0], nullptr, d_y, stream_);
  if (precision == ICARAT_HIP_FP64) {
    launch_apply_k64(L_, kc_, d_young, d_free, d_y, d_q, stream_);
  } else if (precision == ICARAT_HIP_FP64_MFMA) {
    launch_apply_k64_mfma(L_, kc_, d_young, d_free, d_y, d_q, stream_);
  } else if (precision == ICARAT_HIP_FP32 || precision == ICARAT_HIP_BF16) {
    launch_apply_k_mixed(L_, precision == ICARAT_HIP_BF16 ? 1 : 0, d_ke32, d_young32, d_free, d_y,
                         d_q, stream_);
  } else {
    throw Error(kInvalidArgument, "unknown precision");
  }
  launch_lattice_to_numeq(L_, d_dof, numeq, d_q, d_numeq[1], stream_);
  d2h(y, d_numeq[1].p, numeq, stream_);
  sync();
}

void Context::get_diag(double *d) {
  require_density();
  const int numeq = prob_.numeq();
  launch_diag_inv(L_, d_ke, d_young, d_free, d_dinv, d_y, stream_);
  launch_lattice_to_numeq(L_, d_dof, numeq, d_y, d_numeq[1], stream_);
  d2h(d, d_numeq[1].p, numeq, stream_);
  sync();
}

void Context::enqueue_solver_batch(int par) {
  if (batch_precond_ == ICARAT_HIP_PRECOND_MG)
    mg_->enqueue_batch(par);
  else
    enqueue_pcg_batch(-1, par);
}

void Context::build_solver_graph() {
  if (batch_precond_ == ICARAT_HIP_PRECOND_MG)
    mg_->build_graph();
  else
    build_pcg_graph(-1);
}

void Context::launch_solver_batch(int par) {
  build_solver_graph();
  if (batch_precond_ == ICARAT_HIP_PRECOND_MG)
    mg_->launch_batch(par);
  else
    launch_pcg_batch(-1, par);
}

void Context::enqueue_pcg_batch(int kmode, int par) {
  if (kmode < 0) {
    // FP64: the finalizes are folded into the next kernel; one at the end of
    // the batch leaves the scalars complete for the host
    unsigned Gu = 0;
    for (int j = 0; j < kBatch; ++j) {
      double *p_old = d_p[(j + par) & 1], *p_new = d_p[(j + par + 1) & 1];
      const unsigned Gs = launch_pcg_stencil64_fused(L_, kc_, d_young, d_free, d_r, d_dinv, p_old,
                                                     p_new, d_q, d_pcg, d_partials, d_partials2,
                                                     Gu, j, stream_);
      Gu = launch_pcg_update64_fused(n3_, d_x, d_r, p_new, d_q, d_dinv, d_pcg, d_partials2,
                                     d_partials, Gs, j, stream_);
    }
    launch_finalize(kFinPcgBeta, 2, d_partials2, Gu, d_pcg, nullptr, nullptr, stream_);
    return;
  }
  for (int j = 0; j < kBatch; ++j) {
    double *p_old = d_p[(j + par) & 1], *p_new = d_p[(j + par + 1) & 1];
    if (kmode < 0)
      launch_pcg_stencil64(L_, kc_, d_young, d_free, d_r, d_dinv, p_old, p_new, d_q, d_pcg,
                           d_partials, d_counter, stream_);
    else
      launch_pcg_stencil_mfma64(L_, kmode, d_ke32, d_young32, d_free, d_r, d_dinv, p_old, p_new,
                                d_q, d_pcg, d_partials, stream_);
    launch_pcg_update64(n3_, d_x, d_r, p_new, d_q, d_dinv, d_pcg, d_partials, d_counter, stream_);
  }
}

void Context::build_pcg_graph(int kmode) {
  const int m = kmode + 1;
  if (!use_graphs_ || pcg_exec_[m][0]) return;
  for (int par = 0; par < 2; ++par) {
    HIPCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
      enqueue_pcg_batch(kmode, par);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(stream_, &pcg_graph_[m][par]));
    HIPCHK(hipGraphInstantiate(&pcg_exec_[m][par], pcg_graph_[m][par], nullptr, nullptr, 0));
  }
}

void Context::launch_pcg_batch(int kmode, int par) {
  if (use_graphs_)
    HIPCHK(hipGraphLaunch(pcg_exec_[kmode + 1][par], stream_));
  else
    enqueue_pcg_batch(kmode, par);
}

PcgScalars Context::pcg64(double tol, int maxit, bool warm, int kmode, double delta) {
  build_pcg_graph(kmode);
  PcgScalars init{};
  init.tol = tol;
  init.maxit = maxit;
  init.delta = delta;
  *h_pcg.p = init;
  h2d(d_pcg.p, h_pcg.p, 1, stream_);
  if (warm) d2d(d_nv.p, d_x.p, n3_, stream_);
  // x = 0, r = b, p = 0, bnorm2 = b.b
  launch_pcg_init64(n3_, d_b, d_x, d_r, d_p[0], d_dinv, d_pcg, d_partials, d_counter, stream_);
  if (warm) {
    d2d(d_x.p, d_nv.p, n3_, stream_);
    launch_apply_k64(L_, kc_, d_young, d_free, d_x, d_y, stream_);
    launch_replace64(n3_, d_b, d_y, d_r, d_dinv, d_pcg, d_partials, stream_);
  }
  d2h(h_pcg.p, d_pcg.p, 1, stream_);
  sync();
  if (batch_hook_) batch_hook_();
  replacements_ = 0;
  fell_back_ = 0;
  double best = std::sqrt(h_pcg.p->rr);
  int stalls = 0;
  // a low-precision contraction whose recursive residual stagnates never asks
  // for a replacement: force one after kStallWindow iterations
  constexpr int kStallWindow = 1000;
  int last_repl = h_pcg.p->iter;
  while (!h_pcg.p->done || h_pcg.p->status == 0 || h_pcg.p->status == 3) {
    bool forced = false;
    if (!h_pcg.p->done) {
      // iterate until the recursive stop test, a replacement, a cap or a breakdown
      const long long max_batches = (long long)(maxit - h_pcg.p->iter) / kBatch + 3;
      for (long long b = 0; b < max_batches && !h_pcg.p->done; ++b) {
        launch_pcg_batch(kmode, h_pcg.p->iter & 1);
        d2h(h_pcg.p, d_pcg.p, 1, stream_);
        sync();
        if (batch_hook_ && !h_pcg.p->done) batch_hook_();
        if (kmode >= 0 && !h_pcg.p->done && h_pcg.p->iter - last_repl >= kStallWindow) {
          forced = true;
          break;
        }
      }
      if (!h_pcg.p->done && !forced) throw Error(kInternal, "PCG did not reach its stop test");
    }
    if (!forced && h_pcg.p->status != 0 && h_pcg.p->status != 3) break;
    last_repl = h_pcg.p->iter;
    // the FP64 true residual replaces the recursive one (verified stop /
    // reliable update); the search direction is kept
    launch_apply_k64(L_, kc_, d_young, d_free, d_x, d_y, stream_);
    launch_replace64(n3_, d_b, d_y, d_r, d_dinv, d_pcg, d_partials, stream_);
    d2h(h_pcg.p, d_pcg.p, 1, stream_);
    sync();
    if (h_pcg.p->done) break;
    ++replacements_;
    const double rn = std::sqrt(h_pcg.p->rr);
    if (rn < 0.5 * best) {
      best = rn;
      stalls = 0;
    } else if ((++stalls >= 4 || forced) && kmode >= 0) {
      // (a forced replacement that did not halve the residual: the recursive
      // residual stagnated for a whole window)
      // the low-precision contraction stopped making progress: continue the
      // same iteration with the FP64 contraction
      kmode = -1;
      fell_back_ = 1;
      stalls = 0;
      build_pcg_graph(kmode);
      // restart the recurrence (beta = 0): the search directions built with
      // the low-precision operator are not conjugate for the FP64 one
      h_pcg.p->beta = 0.0;
      h2d(d_pcg.p, h_pcg.p, 1, stream_);
    }
    if (h_pcg.p->iter > maxit || replacements_ > 100000) {
      h_pcg.p->status = 1;
      break;
    }
  }
  return *h_pcg.p;
}

void Context::assemble_nodal_solution() {
  d2d(d_u.p, d_x.p, n3_, stream_);
  launch_axpby(n3_, 1.0, d_dirlat, 1.0, d_u, stream_);
}

double Context::true_residual_device() {
  launch_apply_k64(L_, kc_, d_young, d_free, d_u, d_y, stream_);
  launch_residual_norms(n3_, d_f, d_y, d_scal, d_partials, d_counter, stream_);
  d2h(h_scal.p, d_scal.p, 2, stream_);
  sync();
  if (h_scal.p[1] == 0.0) return 0.0;
  return std::sqrt(h_scal.p[0]) / std::sqrt(h_scal.p[1]);
}

double Context::true_residual_host(const double *u) {
  require_density();
  ensure_fext();
  const int numeq = prob_.numeq();
  h2d(d_numeq[0].p, u, numeq, stream_);
  launch_numeq_to_lattice(L_, d_dof, numeq, d_numeq[0], d_dirval, d_u, stream_);
  return true_residual_device();
}

void Context::solve(const icarat_solve_opts &o, double *u_host, icarat_solve_stats *st) {
  require_density();
  if (!(o.tol > 0.0) || o.maxit < 0) throw Error(kInvalidArgument, "tol must be > 0 and maxit >= 0");
  ensure_fext();
  sync();
  const double t0 = now_s();
  prepare_rhs(stream_);
  icarat_solve_stats S{};
  S.status = ICARAT_HIP_OK;
  if (o.precision != ICARAT_HIP_FP64 && o.precision != ICARAT_HIP_FP32 &&
      o.precision != ICARAT_HIP_BF16)
    throw Error(kInvalidArgument, "unknown precision");
  if (o.precond == ICARAT_HIP_PRECOND_JACOBI) {
    const int kmode = o.precision == ICARAT_HIP_FP64 ? -1 : (o.precision == ICARAT_HIP_FP32 ? 0 : 1);
    const double delta = kmode < 0 ? 0.0 : (o.inner_tol > 0.0 && o.inner_tol < 1.0 ? o.inner_tol : 0.1);
    const bool warm = o.warm && kmode < 0 && load_warm_guess(stream_);
    PcgScalars r = pcg64(o.tol, o.maxit, warm, kmode, delta);
    S.iters = r.iter;
    S.refine_steps = replacements_;
    S.fallback = fell_back_;
    S.rec_relres = r.bnorm2 > 0 ? std::sqrt(r.rr) / std::sqrt(r.bnorm2) : 0.0;
    if (r.status == 2) throw Error(kNumerical, "PCG breakdown (non-positive curvature or NaN)");
    if (r.status == 1) S.status = ICARAT_HIP_NOT_CONVERGED;
  } else if (o.precond == ICARAT_HIP_PRECOND_LEARNED) {
    if (!learned_) throw Error(kNotReady, "the learned preconditioner is not set up");
    learned_->solve(o, S);
  } else if (o.precond == ICARAT_HIP_PRECOND_MG) {
    if (!mg_) mg_setup(mg_default_opts());
    mg_->solve(o, S, o.warm && load_warm_guess(stream_));
  } else {
    throw Error(kInvalidArgument, "unknown preconditioner");
  }
  assemble_nodal_solution();
  S.true_relres = true_residual_device();
  save_solution(stream_);
  S.seconds = now_s() - t0;
  copy_out_solution(u_host);
  if (st) *st = S;
}

void Context::copy_out_solution(double *u_host) {
  if (u_host) {
    const int numeq = prob_.numeq();
    launch_lattice_to_numeq(L_, d_dof, numeq, d_u, d_numeq[0], stream_);
    d2h(u_host, d_numeq[0].p, numeq, stream_);
    sync();
  }
  solved_ = true;
}

void Context::prepare_rhs(hipStream_t s) {
  // right-hand side: fext minus the load of the prescribed displacements
  d2d(d_b.p, d_f.p, n3_, s);
  if (prob_.has_nonzero_dirichlet()) {
    launch_apply_k64(L_, kc_, d_young, d_free, d_dirlat, d_y, s);
    launch_axpby(n3_, -1.0, d_y, 1.0, d_b, s);
  }
}

// ---- pieces of the FP64 Jacobi-PCG driver for batched solves -----------------
void Context::pcg_start(double tol, int maxit, hipStream_t s, bool warm) {
  require_density();
  ensure_fext();
  // work on this context's stream (load-case switch, density) precedes stream s
  sync();
  if (!(tol > 0.0) || maxit < 0) throw Error(kInvalidArgument, "tol must be > 0 and maxit >= 0");
  prepare_rhs(s);
  replacements_ = 0;
  fell_back_ = 0;
  if (batch_precond_ == ICARAT_HIP_PRECOND_MG) {
    if (!mg_) mg_setup(mg_default_opts());
    mg_->set_fine_fp32(false);  // batched multigrid solves run the FP64 V-cycle
    mg_->start(tol, maxit, s, warm && load_warm_guess(s));
    return;
  }
  if (batch_precond_ != ICARAT_HIP_PRECOND_JACOBI)
    throw Error(kInvalidArgument, "interleaved solves run Jacobi or multigrid PCG");
  PcgScalars init{};
  init.tol = tol;
  init.maxit = maxit;
  *h_pcg.p = init;
  h2d(d_pcg.p, h_pcg.p, 1, s);
  const bool guess = warm && load_warm_guess(s);
  if (guess) d2d(d_nv.p, d_x.p, n3_, s);
  // x = 0, r = b, p = 0, bnorm2 = b.b; with a guess, r = b - K x replaces r
  launch_pcg_init64(n3_, d_b, d_x, d_r, d_p[0], d_dinv, d_pcg, d_partials, d_counter, s);
  if (guess) {
    d2d(d_x.p, d_nv.p, n3_, s);
    launch_apply_k64(L_, kc_, d_young, d_free, d_x, d_y, s);
    launch_replace64(n3_, d_b, d_y, d_r, d_dinv, d_pcg, d_partials, s);
  }
  d2h(h_pcg.p, d_pcg.p, 1, s);
}

void Context::pcg_verify(hipStream_t s) {
  if (batch_precond_ == ICARAT_HIP_PRECOND_MG) {
    mg_->verify(s);
    ++replacements_;
    return;
  }
  // the FP64 true residual replaces the recursive one; CG continues if it fails the test
  launch_apply_k64(L_, kc_, d_young, d_free, d_x, d_y, s);
  launch_replace64(n3_, d_b, d_y, d_r, d_dinv, d_pcg, d_partials, s);
  d2h(h_pcg.p, d_pcg.p, 1, s);
  ++replacements_;
}

bool Context::pcg_recover(hipStream_t s) {
  return batch_precond_ == ICARAT_HIP_PRECOND_MG && mg_ && mg_->recover_w(s);
}

bool Context::pcg_after_verify() {
  if (batch_precond_ == ICARAT_HIP_PRECOND_MG) return mg_->after_verify();
  return h_pcg.p->done != 0;
}

void Context::pcg_finish(icarat_solve_stats &S) {
  if (batch_precond_ == ICARAT_HIP_PRECOND_MG) mg_->finish();
  const PcgScalars &r = *h_pcg.p;
  S = icarat_solve_stats{};
  S.iters = r.iter;
  S.refine_steps = replacements_;
  S.rec_relres = r.bnorm2 > 0 ? std::sqrt(r.rr) / std::sqrt(r.bnorm2) : 0.0;
  S.status = r.done && r.status == 0 ? ICARAT_HIP_OK : ICARAT_HIP_NOT_CONVERGED;
  if (r.status == 2) throw Error(kNumerical, "PCG breakdown (non-positive curvature or NaN)");
  assemble_nodal_solution();
  S.true_relres = true_residual_device();
  save_solution(stream_);
  solved_ = true;
}

// ---- every load case in one multi-column Jacobi-PCG ---------------------------
bool Context::cases_together(const icarat_solve_opts &o) const {
  if (o.precision != ICARAT_HIP_FP64 || kc_.kernel != 0) return false;
  if (o.precond == ICARAT_HIP_PRECOND_JACOBI) return true;
  return o.precond == ICARAT_HIP_PRECOND_MG && mg_ && mg_->cases_capable(o);
}

void Context::mc_drop_graph() {
  if (mc_.exec) (void)hipGraphExecDestroy(mc_.exec);
  if (mc_.graph) (void)hipGraphDestroy(mc_.graph);
  mc_.exec = nullptr;
  mc_.graph = nullptr;
}

void Context::mc_alloc(int nc) {
  if (mc_.nc == nc) return;
  mc_drop_graph();
  const size_t n = (size_t)nc * n3_;
  mc_.X.alloc(n);
  mc_.R.alloc(n);
  mc_.Q.alloc(n);
  mc_.B.alloc(n);
  mc_.P.alloc(2 * n);
  mc_.pst.alloc((size_t)nc * stencil_blocks(L_));
  mc_.pup.alloc((size_t)nc * 2 * pcg_update_blocks());
  mc_.S.alloc(nc);
  mc_.hS.alloc(nc);
  // the lattice padding (Np rounds the node count up) is read by the vector
  // kernels and never written: it must hold zeros, as in the single-case buffers
  for (DevBuf<double> *b : {&mc_.X, &mc_.R, &mc_.Q, &mc_.B, &mc_.P})
    HIPCHK(hipMemsetAsync(b->p, 0, b->bytes(), stream_));
  sync();
  mc_.nc = nc;
}

void Context::mc_enqueue_batch() {
  const int nc = mc_.nc;
  unsigned Gu = 0;
  for (int j = 0; j < kBatch; ++j) {
    const unsigned Gs = launch_pcg_stencil64_mc(L_, kc_, d_young, d_free, mc_.R, d_dinv, mc_.P,
                                                mc_.Q, mc_.S, mc_.pst, mc_.pup, Gu, j, nc, n3_,
                                                stream_);
    Gu = launch_pcg_update64_mc(n3_, mc_.X, mc_.R, mc_.P, mc_.Q, d_dinv, mc_.S, mc_.pup, mc_.pst,
                                Gs, j, nc, n3_, stream_);
  }
  // each case's scalars complete for the host (iteration count, beta, stop test)
  for (int c = 0; c < nc; ++c)
    launch_finalize(kFinPcgBeta, 2, mc_.pup.p + (size_t)2 * Gu * c, Gu, mc_.S.p + c, nullptr,
                    nullptr, stream_);
}

void Context::mc_prepare() {
  const int nc = prob_.num_load_cases();
  mc_alloc(nc);
  // right-hand sides: each case's fext minus the load of the prescribed displacements
  const bool dir = prob_.has_nonzero_dirichlet();
  if (dir) launch_apply_k64(L_, kc_, d_young, d_free, d_dirlat, d_y, stream_);
  for (int c = 0; c < nc; ++c) {
    double *b = mc_.B.p + (size_t)c * n3_;
    d2d(b, d_fcases_[c].p, n3_, stream_);
    if (dir) launch_axpby(n3_, -1.0, d_y, 1.0, b, stream_);
  }
}

void Context::mc_solve(const icarat_solve_opts &o, icarat_solve_stats *st) {
  if (o.precond == ICARAT_HIP_PRECOND_MG) {
    cases_start(o);
    while (!cases_poll()) {
      cases_launch();
      sync();
    }
    cases_finish(o, st);
    return;
  }
  mc_prepare();
  const int nc = mc_.nc;
  const long long n3 = n3_;
  auto vec = [&](DevBuf<double> &b, int c) { return b.p + (size_t)c * n3; };
  for (int c = 0; c < nc; ++c) {
    PcgScalars init{};
    init.tol = o.tol;
    init.maxit = o.maxit;
    mc_.hS.p[c] = init;
  }
  h2d(mc_.S.p, mc_.hS.p, nc, stream_);
  for (int c = 0; c < nc; ++c) {
    double *x = vec(mc_.X, c), *r = vec(mc_.R, c), *b = vec(mc_.B, c);
    launch_pcg_init64(n3, b, x, r, mc_.P.p + (size_t)2 * c * n3, d_dinv, mc_.S.p + c, d_partials,
                      d_counter, stream_);
    // warm start from the case's last solution: r = b - K x replaces r
    if (o.warm && c < (int)xcase_valid_.size() && xcase_valid_[c]) {
      d2d(x, d_xcases_[c].p, n3, stream_);
      launch_apply_k64(L_, kc_, d_young, d_free, x, d_y, stream_);
      launch_replace64(n3, b, d_y, r, d_dinv, mc_.S.p + c, d_partials, stream_);
    }
  }
  d2h(mc_.hS.p, mc_.S.p, nc, stream_);
  sync();
  if (use_graphs_ && !mc_.exec) {
    HIPCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
      mc_enqueue_batch();
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(stream_, &mc_.graph));
    HIPCHK(hipGraphInstantiate(&mc_.exec, mc_.graph, nullptr, nullptr, 0));
  }
  std::vector<char> fin(nc, 0);
  std::vector<int> repl(nc, 0);
  long long batches = 0;
  const long long max_batches = ((long long)o.maxit / kBatch + 3) * nc + 16;
  for (;;) {
    bool running = false;
    for (int c = 0; c < nc; ++c) {
      if (fin[c]) continue;
      PcgScalars &h = mc_.hS.p[c];
      if (h.done) {
        if (h.status != 0 && h.status != 3) {  // capped or broken down
          fin[c] = 1;
          continue;
        }
        // the FP64 true residual replaces the recursive one (verified stop);
        // the search direction is kept
        launch_apply_k64(L_, kc_, d_young, d_free, vec(mc_.X, c), d_y, stream_);
        launch_replace64(n3, vec(mc_.B, c), d_y, vec(mc_.R, c), d_dinv, mc_.S.p + c, d_partials,
                         stream_);
        d2h(mc_.hS.p + c, mc_.S.p + c, 1, stream_);
        sync();
        if (h.done) {
          fin[c] = 1;
          continue;
        }
        if (h.iter > o.maxit || ++repl[c] > 100000) {
          h.status = 1;
          fin[c] = 1;
          continue;
        }
      }
      running = true;
    }
    if (!running) break;
    if (++batches > max_batches) throw Error(kInternal, "PCG did not reach its stop test");
    if (use_graphs_)
      HIPCHK(hipGraphLaunch(mc_.exec, stream_));
    else
      mc_enqueue_batch();
    d2h(mc_.hS.p, mc_.S.p, nc, stream_);
    sync();
  }
  for (int c = 0; c < nc; ++c) {
    const PcgScalars &r = mc_.hS.p[c];
    if (r.status == 2) throw Error(kNumerical, "PCG breakdown (non-positive curvature or NaN)");
    icarat_solve_stats S{};
    S.status = r.status == 1 ? ICARAT_HIP_NOT_CONVERGED : ICARAT_HIP_OK;
    S.iters = r.iter;
    S.refine_steps = repl[c];
    S.rec_relres = r.bnorm2 > 0 ? std::sqrt(r.rr) / std::sqrt(r.bnorm2) : 0.0;
    if (st) st[c] = S;
  }
  mc_save_warm();
}

void Context::mc_save_warm() {
  // each case's warm start
  for (int c = 0; c < mc_.nc; ++c) {
    if (c >= (int)d_xcases_.size()) {
      d_xcases_.resize(c + 1);
      xcase_valid_.resize(c + 1, 0);
    }
    if (!d_xcases_[c].p) d_xcases_[c].alloc(n3_);
    d2d(d_xcases_[c].p, mc_.X.p + (size_t)c * n3_, n3_, stream_);
    xcase_valid_[c] = 1;
  }
}

// The multigrid solve of every case together, in pieces (the batched engine
// interleaves them with other problems' launches): a case whose W-cycle broke
// down is solved on its own at the end, as its own solve would have gone on.
void Context::cases_start(const icarat_solve_opts &o) {
  if (!mg_) mg_setup(mg_default_opts());
  mc_prepare();
  const int nc = mc_.nc;
  std::vector<char> guess(nc, 0);
  for (int c = 0; c < nc; ++c)
    if (o.warm && c < (int)xcase_valid_.size() && xcase_valid_[c]) {
      d2d(mc_.X.p + (size_t)c * n3_, d_xcases_[c].p, n3_, stream_);
      guess[c] = 1;
    }
  mg_->mc_begin(o, nc, mc_.B, mc_.X, guess.data());
}

bool Context::cases_poll() { return mg_->mc_poll(); }
bool Context::cases_ready() const { return mg_ && mg_->mc_ready(); }
void Context::cases_build_graph() { mg_->mc_build_graph(); }
void Context::cases_launch() { mg_->mc_launch(); }

void Context::cases_finish(const icarat_solve_opts &o, icarat_solve_stats *st) {
  const int nc = mc_.nc;
  std::vector<icarat_solve_stats> S(nc);
  std::vector<char> redo;
  mg_->mc_end(S.data(), redo);
  for (int c = 0; c < nc; ++c) {
    if (!redo[c]) continue;
    use_load_case(c);
    solve(o, nullptr, &S[c]);
    d2d(mc_.X.p + (size_t)c * n3_, d_x.p, n3_, stream_);
  }
  if (st) std::copy(S.begin(), S.end(), st);
  mc_save_warm();
}

void Context::mc_take_case(int c) {
  d2d(d_x.p, mc_.X.p + (size_t)c * n3_, n3_, stream_);
  assemble_nodal_solution();
  d2d(d_f.p, d_fcases_[c].p, n3_, stream_);
  active_case_ = c;
}

void Context::solve_cases(const icarat_solve_opts &o, double *u_host, icarat_solve_stats *st) {
  require_density();
  if (!(o.tol > 0.0) || o.maxit < 0) throw Error(kInvalidArgument, "tol must be > 0 and maxit >= 0");
  if (o.precision != ICARAT_HIP_FP64 && o.precision != ICARAT_HIP_FP32 &&
      o.precision != ICARAT_HIP_BF16)
    throw Error(kInvalidArgument, "unknown precision");
  ensure_fext();
  sync();
  const int nc = prob_.num_load_cases();
  const int numeq = prob_.numeq();
  const int active = prob_.load_case();
  std::vector<icarat_solve_stats> S(nc);
  if (o.precond == ICARAT_HIP_PRECOND_MG && !mg_) mg_setup(mg_default_opts());
  if (!cases_together(o)) {
    // preconditioners and precisions without a multi-column iteration: one case after another
    for (int c = 0; c < nc; ++c) {
      use_load_case(c);
      solve(o, u_host ? u_host + (size_t)c * numeq : nullptr, &S[c]);
    }
    use_load_case(active);
    if (load_warm_guess(stream_)) assemble_nodal_solution();
    solved_ = true;
    if (st) std::copy(S.begin(), S.end(), st);
    return;
  }
  const double t0 = now_s();
  mc_solve(o, S.data());
  for (int c = 0; c < nc; ++c) {
    mc_take_case(c);
    S[c].true_relres = true_residual_device();
    if (u_host) {
      launch_lattice_to_numeq(L_, d_dof, numeq, d_u, d_numeq[0], stream_);
      d2h(u_host + (size_t)c * numeq, d_numeq[0].p, numeq, stream_);
    }
  }
  mc_take_case(active);
  sync();
  const double dt = now_s() - t0;
  for (auto &s : S) s.seconds = dt;
  solved_ = true;
  if (st) std::copy(S.begin(), S.end(), st);
}

void Context::simp_cases_together(const icarat_solve_opts &so) {
  const int nc = prob_.num_load_cases();
  std::vector<icarat_solve_stats> S(nc);
  const double t0 = now_s();
  mc_solve(so, S.data());
  simp_.S.solve_s += now_s() - t0;
  simp_cases_post(S.data());
}

void Context::simp_cases_post(icarat_solve_stats *S) {
  const int nc = prob_.num_load_cases();
  for (int lc = 0; lc < nc; ++lc) {
    const double t1 = now_s();
    mc_take_case(lc);
    S[lc].true_relres = true_residual_device();
    solved_ = true;
    simp_.S.solve_s += now_s() - t1;
    simp_case_post(lc, S[lc]);
  }
}

void Context::compliance_sensitivity(double vnorm, double *c, double *dfds, double *dgds) {
  require_density();
  ensure_fext();
  if (!solved_) throw Error(kNotReady, "no solution for the current density");
  if (!(vnorm != 0.0)) throw Error(kInvalidArgument, "vnorm must be non-zero");
  const auto &g = prob_.spec();
  if (c) {
    launch_dot(n3_, d_f, d_u, d_scal, d_partials, d_counter, stream_);
    d2h(h_scal.p, d_scal.p, 1, stream_);
  }
  if (dfds || dgds) {
    launch_flush_small(n3_, d_u, d_nv, stream_);
    launch_sensitivity64(L_, d_ke, d_dens, d_nv, g.young1, g.young2, g.penalty, vol_e_, vnorm,
                         d_e[0], d_e[1], stream_);
    if (dfds) d2h(dfds, d_e[0].p, L_.nelx, stream_);
    if (dgds) d2h(dgds, d_e[1].p, L_.nelx, stream_);
  }
  sync();
  if (c) *c = h_scal.p[0];
}

void Context::mises_host(double *mises, double *stress) {
  require_density();
  ensure_fext();
  if (!solved_) throw Error(kNotReady, "no solution for the current density");
  const auto &g = prob_.spec();
  const double h[3] = {g.lx / g.nx, g.ly / g.ny, g.lz / g.nz};
  DevBuf<double> d_st;
  if (stress) d_st.alloc(6 * (size_t)L_.nelx);
  launch_flush_small(n3_, d_u, d_nv, stream_);
  launch_mises64(L_, d_dens, d_nv, g.young1, g.young2, g.penalty, g.poisson, h, d_e[6],
                 stress ? d_st.p : nullptr, stream_);
  d2h(mises, d_e[6].p, L_.nelx, stream_);
  if (stress) d2h(stress, d_st.p, 6 * L_.nelx, stream_);
  sync();
}

void Context::nodal_values_host(double *u) {
  if (!solved_) throw Error(kNotReady, "no solution for the current density");
  launch_flush_small(n3_, d_u, d_nv, stream_);
  std::vector<double> lat(n3_);
  d2h(lat.data(), d_nv.p, n3_, stream_);
  sync();
  for (long long n = 0; n < L_.numnp; ++n)
    for (int c = 0; c < 3; ++c) u[3 * n + c] = lat[c * L_.Np + n];
}

// ---- filter / OC ----------------------------------------------------------
// element work vectors
enum { E_DFDS = 0, E_DGDS = 1, E_AA = 2, E_XF = 3, E_XP = 4, E_THD = 5, E_TMP = 6, E_DES = 7 };

void Context::filter_setup(double radius) {
  ensure_device();
  if (!(radius > 1e-10)) throw Error(kInvalidArgument, "filter radius must be > 1e-10");
  std::vector<int> off;
  std::vector<double> w;
  int R = 0;
  prob_.filter_table(radius, off, w, R);
  if (R > 3) throw Error(kInvalidArgument, "filter radius above 3 element widths is not supported");
  R = std::max(R, 1);
  const int count = (int)w.size();
  std::vector<int> lin(count);
  filter_lds_offsets(off.data(), count, R, lin.data());
  d_flin.alloc(count);
  d_fw.alloc(count);
  h2d(d_flin.p, lin.data(), count, stream_);
  h2d(d_fw.p, w.data(), count, stream_);
  ftab_ = FilterTable{count, R, d_flin.p, d_fw.p};
  // normaliser sum_j w_ij (geometry only), so the chain rule works before a
  // density pass as well
  launch_axpby(L_.nelx, 0.0, d_e[E_TMP], 0.0, d_e[E_TMP], stream_);
  launch_density_filter(L_, ftab_, d_e[E_TMP], d_e[E_AA], d_e[E_XF], 0, stream_);
  sync();
  radius_ = radius;
  filter_ready_ = true;
}

void Context::filter_device(const double *in, double *out, int is_sens) {
  if (!filter_ready_) throw Error(kNotReady, "filter_setup has not been called");
  launch_density_filter(L_, ftab_, in, d_e[E_AA], out, is_sens, stream_);
}

void Context::sensitivity_filter_host(const double *x, const double *dfds, double *out) {
  if (!filter_ready_) throw Error(kNotReady, "filter_setup has not been called");
  const long long n = L_.nelx;
  if (d_sfilt_.n < (size_t)(3 * n)) d_sfilt_.alloc(3 * n);
  h2d(d_sfilt_.p, x, n, stream_);
  h2d(d_sfilt_.p + n, dfds, n, stream_);
  launch_sensitivity_filter(L_, ftab_, d_sfilt_.p, d_sfilt_.p + n, d_sfilt_.p + 2 * n, vol_e_,
                            stream_);
  d2h(out, d_sfilt_.p + 2 * n, n, stream_);
  sync();
}

void Context::filter_host(const double *in, double *out, int is_sens) {
  h2d(d_e[E_TMP].p, in, L_.nelx, stream_);
  filter_device(d_e[E_TMP], d_e[E_XF], is_sens);
  d2h(out, d_e[E_XF].p, L_.nelx, stream_);
  sync();
}

void Context::threshold_host(const double *in, double beta, double T, double *out, int deriv) {
  ensure_device();
  h2d(d_e[E_TMP].p, in, L_.nelx, stream_);
  launch_threshold(L_.nelx, d_e[E_TMP], beta, T, d_e[E_XF], deriv, stream_);
  d2h(out, d_e[E_XF].p, L_.nelx, stream_);
  sync();
}

void Context::helm_setup(double radius) {
  helm_.reset();
  helm_.reset(new HelmFilter(*this, radius));
}

void Context::helm_host(const double *in, double *out, int is_sens) {
  if (!helm_) throw Error(kNotReady, "helm_setup has not been called");
  h2d(d_e[E_TMP].p, in, L_.nelx, stream_);
  helm_->apply(d_e[E_TMP], d_e[E_XF], is_sens);
  d2h(out, d_e[E_XF].p, L_.nelx, stream_);
  sync();
}

void Context::loop_filter(const double *in, double *out, int is_sens) {
  if (simp_.o.filter == 1) helm_->apply(in, out, is_sens);
  else filter_device(in, out, is_sens);
}

int Context::oc_device_helm(double *x, const double *dc, const double *dv, double g, double xmin,
                            double xmax, int use_projection, double beta) {
  if (!helm_) throw Error(kNotReady, "helm_setup has not been called");
  OcState s0{};
  s0.l0 = 0.0;
  s0.l1 = 1.0e4;
  s0.g = g;
  s0.maxiter = 100;
  *h_oc.p = s0;
  h2d(d_oc.p, h_oc.p, 1, stream_);
  // x' at lambda -> E_TMP, filtered -> E_XP (both free while the design is
  // updated), volume of its projection -> bisection update.  Without the
  // projection the filter is skipped: the Helmholtz filter preserves the
  // volume exactly (A 1 = M 1 = (V_e / 8) S^T 1 with no boundary condition, so
  // sum_e (W x)_e = sum_e x_e), and the volume of x' is the volume of its
  // filtered field to rounding.
  constexpr int kSteps = 4;
  for (int b = 0; b < 100 / kSteps + 2; ++b) {
    for (int k = 0; k < kSteps; ++k) {
      launch_oc_candidate(L_.nelx, x, dc, dv, xmin, xmax, d_oc, d_e[E_TMP], stream_);
      const double *xf = d_e[E_TMP];
      if (use_projection) {
        helm_->apply(d_e[E_TMP], d_e[E_XP], 0);
        xf = d_e[E_XP];
      }
      launch_oc_volume(L_.nelx, xf, use_projection, beta, vol_e_, d_oc, d_partials, stream_);
    }
    d2h(h_oc.p, d_oc.p, 1, stream_);
    sync();
    if (h_oc.p->done) break;
  }
  if (!h_oc.p->done) throw Error(kInternal, "OC bisection did not stop");
  launch_oc_final(L_.nelx, x, dc, dv, xmin, xmax, d_oc, stream_);
  return h_oc.p->counter;
}

int Context::oc_device(double *x, const double *dc, const double *dv, double g, double xmin,
                       double xmax, int use_projection, double beta) {
  if (!filter_ready_) throw Error(kNotReady, "filter_setup has not been called");
  OcState s0{};
  s0.l0 = 0.0;
  s0.l1 = 1.0e4;
  s0.g = g;
  s0.maxiter = 100;
  *h_oc.p = s0;
  h2d(d_oc.p, h_oc.p, 1, stream_);
  constexpr int kSteps = 8;
  for (int b = 0; b < 100 / kSteps + 2; ++b) {
    for (int k = 0; k < kSteps; ++k)
      launch_oc_step(L_, ftab_, x, dc, dv, xmin, xmax, use_projection, beta, vol_e_, d_oc,
                     d_partials, d_counter, stream_);
    d2h(h_oc.p, d_oc.p, 1, stream_);
    sync();
    if (h_oc.p->done) break;
  }
  if (!h_oc.p->done) throw Error(kInternal, "OC bisection did not stop");
  launch_oc_final(L_.nelx, x, dc, dv, xmin, xmax, d_oc, stream_);
  return h_oc.p->counter;
}

int Context::oc_host(double *x, const double *dfdx, double g, const double *dgdx, double xmin,
                     double xmax, int use_projection, double beta) {
  h2d(d_e[E_DES].p, x, L_.nelx, stream_);
  h2d(d_e[E_DFDS].p, dfdx, L_.nelx, stream_);
  h2d(d_e[E_DGDS].p, dgdx, L_.nelx, stream_);
  const int steps = oc_device(d_e[E_DES], d_e[E_DFDS], d_e[E_DGDS], g, xmin, xmax,
                              use_projection, beta);
  d2h(x, d_e[E_DES].p, L_.nelx, stream_);
  sync();
  return steps;
}

// ---- the design loop ----------------------------------------------------
// ---- the SIMP loop, in phases (shared by simp_run and the batch driver) ------
void Context::simp_begin(const icarat_simp_opts &o, const double *x) {
  require_dof();
  ensure_fext();
  if (o.filter != 0 && o.filter != 1)
    throw Error(kInvalidArgument, "filter must be 0 (cone) or 1 (Helmholtz)");
  if (o.filter == 0 && !filter_ready_) throw Error(kNotReady, "filter_setup has not been called");
  if (o.filter == 1 && !helm_) throw Error(kNotReady, "helm_setup has not been called");
  if (o.n_iter < 0) throw Error(kInvalidArgument, "n_iter must be >= 0");
  if (!(o.conv_eps >= 0.0)) throw Error(kInvalidArgument, "conv_eps must be >= 0");
  if (o.optimizer != 0 && o.optimizer != 1 && o.optimizer != 2)
    throw Error(kInvalidArgument, "optimizer must be 0 (OC), 1 (MMA) or 2 (GCMMA)");
  sync();
  simp_ = SimpState{};
  simp_.o = o;
  simp_.t_all = now_s();
  h2d(d_e[E_DES].p, x, L_.nelx, stream_);
  if (o.optimizer == 1) {
    if (!mma_) mma_.reset(new DeviceMMA(*this));
    mma_->reset();
  }
}

void Context::simp_pre(int step) {
  const icarat_simp_opts &o = simp_.o;
  const long long n = L_.nelx;
  double t0 = now_s();
  // x_tilde = filter(x); x_hat = projection(x_tilde) or x_tilde
  loop_filter(d_e[E_DES], d_e[E_XF], 0);
  simp_.beta = std::min(std::pow(2, (int)(step / 40)), 16.0);
  double *xp = d_e[E_XF];
  if (o.use_projection) {
    launch_threshold(n, d_e[E_XF], simp_.beta, 0.5, d_e[E_XP], 0, stream_);
    launch_threshold(n, d_e[E_XF], simp_.beta, 0.5, d_e[E_THD], 1, stream_);
    xp = d_e[E_XP];
  }
  d2d(d_dens.p, xp, n, stream_);
  set_density_device();
  // volume of the physical design (V0 of the first iteration normalises MMA's constraint)
  launch_volume(n, d_dens, vol_e_, d_scal.p + 1, d_partials, d_counter, stream_);
  d2h(h_scal.p + 1, d_scal.p + 1, 1, stream_);
  sync();
  simp_.vol = h_scal.p[1];
  if (step == 1) simp_.V0 = simp_.vol;
  simp_.c = 0.0;
  simp_.relres = 0.0;
  simp_.iters = 0;
  simp_.S.filter_s += now_s() - t0;
}

void Context::simp_case_post(int lc, const icarat_solve_stats &ss) {
  const icarat_simp_opts &o = simp_.o;
  const auto &g = prob_.spec();
  const long long n = L_.nelx;
  const double t0 = now_s();
  simp_.iters += ss.iters;
  simp_.relres = std::max(simp_.relres, ss.true_relres);
  simp_.S.pcg_iters += ss.iters;
  simp_.S.fallbacks += ss.fallback;
  simp_.S.max_true_relres = std::max(simp_.S.max_true_relres, ss.true_relres);
  launch_dot(n3_, d_f, d_u, d_scal, d_partials, d_counter, stream_);
  launch_flush_small(n3_, d_u, d_nv, stream_);
  // dgds does not depend on u: the first case's is kept
  launch_sensitivity64(L_, d_ke, d_dens, d_nv, g.young1, g.young2, g.penalty, vol_e_,
                       o.optimizer == 1 ? simp_.V0 : 1.0, lc == 0 ? d_e[E_DFDS] : d_e[E_TMP],
                       lc == 0 ? d_e[E_DGDS] : d_e[E_XP], stream_);
  if (lc > 0) launch_axpby(n, 1.0, d_e[E_TMP], 1.0, d_e[E_DFDS], stream_);
  d2h(h_scal.p, d_scal.p, 1, stream_);
  sync();
  simp_.c += h_scal.p[0];
  simp_.S.sens_s += now_s() - t0;
}

void Context::simp_update(int step, double *hist) {
  const icarat_simp_opts &o = simp_.o;
  const long long n = L_.nelx;
  double t2 = now_s();
  if (o.use_projection) {
    launch_mul(n, d_e[E_THD], d_e[E_DFDS], stream_);
    launch_mul(n, d_e[E_THD], d_e[E_DGDS], stream_);
  }
  loop_filter(d_e[E_DFDS], d_e[E_TMP], 1);
  std::swap(d_e[E_DFDS], d_e[E_TMP]);
  loop_filter(d_e[E_DGDS], d_e[E_TMP], 1);
  std::swap(d_e[E_DGDS], d_e[E_TMP]);
  sync();
  double t3 = now_s();
  simp_.S.filter_s += t3 - t2;
  if (hist) {
    hist[4 * (step - 1) + 0] = simp_.c;
    hist[4 * (step - 1) + 1] = simp_.vol;
    hist[4 * (step - 1) + 2] = (double)simp_.iters;
    hist[4 * (step - 1) + 3] = simp_.relres;
  }
  // the drivers' history line and convergence test, both after the compliance
  // (opt_macro_oc.cpp:134-137, optimization/opt_base.hpp:28-44); the design
  // update of the converged iteration still runs
  if (o.history_csv) add_plot(o.history_csv, (double)step, simp_.c, step == 1);
  if (o.conv_eps > 0.0) {
    if (step > o.conv_minstep && std::fabs(simp_.c_old - simp_.c) / simp_.c < o.conv_eps)
      simp_.S.converged = 1;
    else
      simp_.c_old = simp_.c;
  }
  simp_.S.steps = step;
  if (o.optimizer == 1) {
    const double gc = step == 1 ? 0.0 : simp_.vol / simp_.V0 - 1.0;
    mma_->update(d_e[E_DES], d_e[E_DFDS], gc, d_e[E_DGDS], o.xmin, o.xmax);
  } else if (o.filter == 1) {
    oc_device_helm(d_e[E_DES], d_e[E_DFDS], d_e[E_DGDS], simp_.V0, o.xmin, o.xmax,
                   o.use_projection, simp_.beta);
  } else {
    oc_device(d_e[E_DES], d_e[E_DFDS], d_e[E_DGDS], simp_.V0, o.xmin, o.xmax, o.use_projection,
              simp_.beta);
  }
  sync();
  simp_.S.update_s += now_s() - t3;
}

void Context::simp_end(double *x, icarat_simp_stats *st) {
  d2h(x, d_e[E_DES].p, L_.nelx, stream_);
  sync();
  simp_.S.seconds = now_s() - simp_.t_all;
  if (st) *st = simp_.S;
}

DeviceGCMMA &Context::gcmma() {
  ensure_device();
  if (!gcmma_) gcmma_.reset(new DeviceGCMMA(*this));
  return *gcmma_;
}

void Context::gcmma_host(int mode, double *xmma, double f0xnew, double fxnew, const double *xval,
                         double f0x, const double *df0dx, double fx, const double *dfdx,
                         double xmin, double xmax) {
  DeviceGCMMA &G = gcmma();
  const long long n = L_.nelx;
  h2d(d_e[E_DES].p, xval, n, stream_);
  h2d(d_e[E_DFDS].p, df0dx, n, stream_);
  h2d(d_e[E_DGDS].p, dfdx, n, stream_);
  if (mode == 0) {
    G.outer(d_e[E_XF], d_e[E_DES], f0x, d_e[E_DFDS], fx, d_e[E_DGDS], xmin, xmax);
  } else {
    h2d(d_e[E_XF].p, xmma, n, stream_);
    G.inner(d_e[E_XF], f0xnew, fxnew, d_e[E_DES], f0x, d_e[E_DFDS], fx, d_e[E_DGDS], xmin, xmax);
  }
  d2h(xmma, d_e[E_XF].p, n, stream_);
  sync();
}

double Context::gcmma_evaluate(const double *xd, bool sens, double *vol) {
  const icarat_simp_opts &o = simp_.o;
  const long long n = L_.nelx;
  double t0 = now_s();
  loop_filter(xd, d_e[E_XF], 0);
  double *xp = d_e[E_XF];
  if (o.use_projection) {
    launch_threshold(n, d_e[E_XF], simp_.beta, 0.5, d_e[E_XP], 0, stream_);
    launch_threshold(n, d_e[E_XF], simp_.beta, 0.5, d_e[E_THD], 1, stream_);
    xp = d_e[E_XP];
  }
  d2d(d_dens.p, xp, n, stream_);
  set_density_device();
  launch_volume(n, d_dens, vol_e_, d_scal.p + 1, d_partials, d_counter, stream_);
  d2h(h_scal.p + 1, d_scal.p + 1, 1, stream_);
  sync();
  *vol = h_scal.p[1];
  simp_.S.filter_s += now_s() - t0;
  double c = 0.0;
  const int ncase = prob_.num_load_cases();
  for (int lc = 0; lc < ncase; ++lc) {
    if (ncase > 1) use_load_case(lc);
    const double t1 = now_s();
    icarat_solve_stats ss{};
    solve(o.solve, nullptr, &ss);
    simp_.S.solve_s += now_s() - t1;
    if (sens) {
      simp_.c = 0.0;
      simp_case_post(lc, ss);
      c += simp_.c;
    } else {
      simp_.iters += ss.iters;
      simp_.relres = std::max(simp_.relres, ss.true_relres);
      simp_.S.pcg_iters += ss.iters;
      simp_.S.max_true_relres = std::max(simp_.S.max_true_relres, ss.true_relres);
      launch_dot(n3_, d_f, d_u, d_scal, d_partials, d_counter, stream_);
      d2h(h_scal.p, d_scal.p, 1, stream_);
      sync();
      c += h_scal.p[0];
    }
  }
  if (ncase > 1) use_load_case(prob_.load_case());
  if (sens) {
    const double t2 = now_s();
    if (o.use_projection) {
      launch_mul(n, d_e[E_THD], d_e[E_DFDS], stream_);
      launch_mul(n, d_e[E_THD], d_e[E_DGDS], stream_);
    }
    loop_filter(d_e[E_DFDS], d_e[E_TMP], 1);
    std::swap(d_e[E_DFDS], d_e[E_TMP]);
    loop_filter(d_e[E_DGDS], d_e[E_TMP], 1);
    std::swap(d_e[E_DGDS], d_e[E_TMP]);
    sync();
    simp_.S.filter_s += now_s() - t2;
  }
  return c;
}

void Context::gcmma_run(const icarat_simp_opts &o, double *x, double *hist,
                        icarat_simp_stats *st) {
  simp_begin(o, x);
  DeviceGCMMA &G = gcmma();
  G.reset();
  DevBuf<double> xnew(L_.nelx);
  for (int step = 1; step <= o.n_iter; ++step) {
    // beta schedule of the GCMMA driver (opt_macro_gcmma.cpp:91-93)
    const int bt = (int)std::max((step - 100) / 40.0, 1.0);
    simp_.beta = std::min((double)bt, 8.0);
    simp_.iters = 0;
    simp_.relres = 0.0;
    double vol = 0.0;
    const double c = gcmma_evaluate(d_e[E_DES], true, &vol);
    if (step == 1) simp_.V0 = vol;
    const double g = vol - simp_.V0;
    double t0 = now_s();
    G.outer(xnew, d_e[E_DES], c, d_e[E_DFDS], g, d_e[E_DGDS], o.xmin, o.xmax);
    simp_.S.update_s += now_s() - t0;
    double vnew = 0.0;
    double cnew = gcmma_evaluate(xnew, false, &vnew);
    double gnew = vnew - simp_.V0;
    bool ok = G.concheck(cnew, gnew);
    for (int inner = 0; !ok && inner < 15; ++inner) {
      t0 = now_s();
      G.inner(xnew, cnew, gnew, d_e[E_DES], c, d_e[E_DFDS], g, d_e[E_DGDS], o.xmin, o.xmax);
      simp_.S.update_s += now_s() - t0;
      cnew = gcmma_evaluate(xnew, false, &vnew);
      gnew = vnew - simp_.V0;
      ok = G.concheck(cnew, gnew);
      simp_.S.inner_iters += 1;
    }
    d2d(d_e[E_DES].p, xnew.p, L_.nelx, stream_);
    if (hist) {
      hist[4 * (step - 1) + 0] = c;
      hist[4 * (step - 1) + 1] = vol;
      hist[4 * (step - 1) + 2] = (double)simp_.iters;
      hist[4 * (step - 1) + 3] = simp_.relres;
    }
  }
  simp_end(x, st);
}

void Context::simp_run(const icarat_simp_opts &o, double *x, double *hist, icarat_simp_stats *st) {
  if (o.optimizer == 2) {
    gcmma_run(o, x, hist, st);
    return;
  }
  simp_begin(o, x);
  for (int step = 1; step <= o.n_iter; ++step) {
    simp_pre(step);
    // every load case: solve, compliance f.u, sensitivity; both summed over the
    // cases (reference opt_multi_load.cpp:118-146)
    const int ncase = prob_.num_load_cases();
    if (ncase > 1 && o.solve.precond == ICARAT_HIP_PRECOND_MG && !mg_) mg_setup(mg_default_opts());
    // automatic (case_mode 0): cases_auto
    const bool together = ncase > 1 && cases_together(o.solve) &&
                          (o.case_mode == 2 || (o.case_mode == 0 && cases_auto(o.solve)));
    if (together) {
      // one multi-column iteration over the cases (the same arithmetic per case)
      simp_cases_together(o.solve);
      use_load_case(prob_.load_case());
      simp_update(step, hist);
      if (simp_.S.converged) break;
      continue;
    }
    for (int lc = 0; lc < ncase; ++lc) {
      if (ncase > 1) use_load_case(lc);
      const double t1 = now_s();
      icarat_solve_stats ss{};
      icarat_solve_opts so = o.solve;
      // the first train_iters solves of the preconditioner's lifetime are
      // solved by Jacobi-PCG and feed the training set
      const bool harvest = so.precond == ICARAT_HIP_PRECOND_LEARNED && learned_ &&
                           learned_->harvested() < learned_->opts().train_iters;
      if (harvest) {
        // the base method's own solve (Jacobi-PCG or multigrid PCG) is harvested
        so.precond = learned_->opts().base == 1 ? ICARAT_HIP_PRECOND_MG : ICARAT_HIP_PRECOND_JACOBI;
        learned_->begin_harvest();
        batch_hook_ = [this]() { learned_->snapshot(); };
        try {
          solve(so, nullptr, &ss);
        } catch (...) {
          batch_hook_ = nullptr;
          learned_->end_harvest(false);
          throw;
        }
        batch_hook_ = nullptr;
        learned_->end_harvest(ss.status == ICARAT_HIP_OK);
        learned_->note_base_iters(ss.iters);
        const int ti = std::max(1, learned_->opts().train_iters);
        const int h = learned_->harvested();  // 1..ti
        const int steps =
            learned_->opts().train_steps / ti + (h <= learned_->opts().train_steps % ti ? 1 : 0);
        const double tt = now_s();
        learned_->train(steps);
        simp_.S.train_s += now_s() - tt;
      } else {
        const bool trial = so.precond == ICARAT_HIP_PRECOND_LEARNED && learned_ &&
                           learned_->trial_pending();
        solve(so, nullptr, &ss);
        if (trial) learned_->trial_result(ss.iters);
      }
      simp_.S.solve_s += now_s() - t1;
      simp_case_post(lc, ss);
    }
    if (ncase > 1) use_load_case(prob_.load_case());
    simp_update(step, hist);
    if (simp_.S.converged) break;
  }
  simp_end(x, st);
}

// ---- learned preconditioner --------------------------------------------------
void Context::learned_setup(const icarat_learned_opts &o) {
  require_dof();
  if (o.train_iters < 0 || o.train_steps < 0 || o.snapshot_every < 1 || o.max_samples < 1 ||
      o.batch_iters < 1 || o.batch_iters > 64 || o.stall_iters < 1 || !(o.lr > 0.f))
    throw Error(kInvalidArgument, "invalid learned-preconditioner options");
  learned_.reset();
  learned_.reset(new LearnedPrecond(*this, o));
}

LearnedPrecond &Context::learned() {
  if (!learned_) throw Error(kNotReady, "the learned preconditioner is not set up");
  return *learned_;
}

void Context::learned_harvest(const icarat_solve_opts &o, double *u_host, icarat_solve_stats *st) {
  LearnedPrecond &lp = learned();
  require_density();
  icarat_solve_opts so = o;
  so.precond = lp.opts().base == 1 ? ICARAT_HIP_PRECOND_MG : ICARAT_HIP_PRECOND_JACOBI;
  lp.begin_harvest();
  batch_hook_ = [&lp]() { lp.snapshot(); };
  icarat_solve_stats ss{};
  try {
    solve(so, u_host, &ss);
  } catch (...) {
    batch_hook_ = nullptr;
    lp.end_harvest(false);
    throw;
  }
  batch_hook_ = nullptr;
  lp.end_harvest(ss.status == ICARAT_HIP_OK);
  if (st) *st = ss;
}

void Context::learned_apply(const double *r, double *z) {
  LearnedPrecond &lp = learned();
  require_density();
  const int numeq = prob_.numeq();
  h2d(d_numeq[0].p, r, numeq, stream_);
  // constrained DOFs are ignored by M (D^-1 is zero there)
  launch_numeq_to_lattice(L_, d_dof, numeq, d_numeq[0], d_dirval, d_y, stream_);
  lp.apply(d_y, d_nv);
  launch_lattice_to_numeq(L_, d_dof, numeq, d_nv, d_numeq[1], stream_);
  d2h(z, d_numeq[1].p, numeq, stream_);
  sync();
}

void Context::drop_graphs() {
  for (int m = 0; m < 3; ++m)
    for (int p = 0; p < 2; ++p) {
      if (pcg_exec_[m][p]) (void)hipGraphExecDestroy(pcg_exec_[m][p]);
      if (pcg_graph_[m][p]) (void)hipGraphDestroy(pcg_graph_[m][p]);
      pcg_exec_[m][p] = nullptr;
      pcg_graph_[m][p] = nullptr;
    }
  mc_drop_graph();
  if (learned_) learned_->drop_graphs();
  if (mg_) mg_->drop_graphs();
}

void Context::mg_setup(const icarat_mg_opts &o) {
  require_dof();
  if (o.max_levels < 1 || o.nu_pre < 0 || o.nu_post < 1 || o.nu_coarse < 1 || !(o.omega > 0.0) ||
      !(o.omega < 2.0) || o.batch_iters < 1 || o.batch_iters > 64 || o.dense_dofs < -1 ||
      o.direct_nodes < 0 || o.coarse_avg < 0 || o.coarse_avg > 5 || o.w_levels < -1 ||
      !(o.w_switch >= 0.0) || o.compensated < 0 || o.compensated > 1)
    throw Error(kInvalidArgument, "invalid multigrid options");
  sync();
  mg_.reset();
  mg_.reset(new Multigrid(*this, o));
}

void Context::set_stencil_tuning(int kernel, int nchunks) {
  ensure_device();
  if (kernel < 0 || kernel > 1 || nchunks < 0) throw Error(kInvalidArgument, "bad stencil tuning");
  sync();
  kc_.kernel = kernel;
  kc_.nchunks = nchunks;
  drop_graphs();
}

void Context::set_stencil(int kernel, int sym_max) {
  ensure_device();
  KeCoef k = make_ke_coef(prob_.ke0(), d_ke.p);
  if (k.sym > sym_max) {
    // fall to the highest allowed class that still reproduces KE0
    k.sym = 0;
    if (sym_max >= 1) {
      KeCoef box = make_ke_coef_class(prob_.ke0(), d_ke.p, 1);
      if (box.sym == 1) k = box;
    }
  }
  k.kernel = kernel;
  k.ke32 = d_ke32.p;
  k.xlo = kc_.xlo;
  sync();
  kc_ = k;
  drop_graphs();
}

// ---- measurement hook ------------------------------------------------------
void Context::time_kernel(int kernel, int reps, double *avg_ms, double *flops, double *bytes) {
  require_density();
  ensure_fext();
  if (reps <= 0) throw Error(kInvalidArgument, "reps must be positive");
  const double N = (double)L_.numnp, E = (double)L_.nelx;
  double F = 0.0, B = 0.0;
  std::function<void()> launch;
  DevBuf<double> hook_xlo;
  if (kernel == 0 || (kernel >= 3 && kernel <= 6)) {
    // p = dinv r + beta p_old with r = f (non-zero) and beta = 0, so p.q > 0
    // and the stop flag never rises between launches
    d2d(d_r.p, d_f.p, n3_, stream_);
    PcgScalars s0{};
    s0.rho = 1.0;
    s0.tol = 0.0;
    s0.maxit = 1 << 30;
    *h_pcg.p = s0;
    h2d(d_pcg.p, h_pcg.p, 1, stream_);
    KeCoef kc = kc_;
    if (kernel == 3) kc.sym = 0;                 // KE0 from memory
    if (kernel == 4) kc.kernel = kc.kernel == 1 ? 0 : 1;  // tiles, or z-marching if tiles are set
    if (kernel == 0) {
      // the fused stencil of the solver (in-batch iteration 1): its beta step
      // reduces update partials; ones keep every stop test false
      std::vector<double> ones(d_partials2.n, 1.0);
      h2d(d_partials2.p, ones.data(), ones.size(), stream_);
      h_pcg.p->rho_slot[0] = h_pcg.p->rho_slot[1] = 1.0;
      h_pcg.p->iter_base = 0;
      h2d(d_pcg.p, h_pcg.p, 1, stream_);
      launch = [this]() {
        launch_pcg_stencil64_fused(L_, kc_, d_young, d_free, d_r, d_dinv, d_p[0], d_p[1], d_q,
                                   d_pcg, d_partials, d_partials2, 512, 1, stream_);
      };
    } else if (kernel <= 4) {
      launch = [this, kc]() {
        launch_pcg_stencil64(L_, kc, d_young, d_free, d_r, d_dinv, d_p[0], d_p[1], d_q, d_pcg,
                             d_partials, d_counter, stream_, false);
      };
    } else {
      const int mode = kernel - 5;  // 0 FP32, 1 BF16 MFMA contraction
      launch = [this, mode]() {
        launch_pcg_stencil_mfma64(L_, mode, d_ke32, d_young32, d_free, d_r, d_dinv, d_p[0], d_p[1],
                                  d_q, d_pcg, d_partials, stream_, false);
      };
    }
    F = 2.0 * 24 * 24 * E;                      // dense KE0 x U_e contraction
    B = 3.0 * N * 8 * 3 + 2.0 * 3 * N * 8 + E * 8 + N;  // r, dinv, p_old in; p, q out; young; mask
  } else if (kernel == 1) {
    // the fused update of the solver: alpha = rho / p.q from stencil partials
    // (ones: p.q = block count > 0), then the vector update
    PcgScalars s0{};
    s0.rho = 1.0;
    s0.tol = 0.0;
    s0.maxit = 1 << 30;
    s0.rho_slot[0] = s0.rho_slot[1] = 0.0;  // alpha = 0: the vectors do not drift
    *h_pcg.p = s0;
    h2d(d_pcg.p, h_pcg.p, 1, stream_);
    const unsigned gs = (unsigned)std::min<long long>(stencil_blocks(L_), 4096);
    std::vector<double> ones(gs, 1.0);
    h2d(d_partials.p, ones.data(), gs, stream_);
    launch = [this, gs]() {
      launch_pcg_update64_fused(n3_, d_x, d_r, d_p[1], d_q, d_dinv, d_pcg, d_partials2,
                                d_partials, gs, 1, stream_);
    };
    F = 8.0 * 3 * N;
    B = (5.0 + 2.0) * 3 * N * 8;                // x, r, p, q, dinv in; x, r out
  } else if (kernel == 2) {
    if (!filter_ready_) throw Error(kNotReady, "filter_setup has not been called");
    launch = [&]() { launch_density_filter(L_, ftab_, d_dens, d_e[2], d_e[3], 0, stream_); };
    F = 4.0 * ftab_.count * E;
    B = 3.0 * 8 * E;                            // x in; x_tilde, aa out
  } else if (kernel == 7) {
    launch = [&]() { launch_copy64(n3_, d_x, d_y, stream_); };
    F = 0.0;
    B = 2.0 * 8 * n3_;                          // x in, y out
  } else if (kernel == 8) {
    // y = K x as the multigrid smoother runs it on the fine lattice
    d2d(d_r.p, d_f.p, n3_, stream_);
    launch = [this]() { launch_apply_k64(L_, kc_, d_young, d_free, d_r, d_y, stream_); };
    F = 2.0 * 24 * 24 * E;
    B = 2.0 * 3 * N * 8 + E * 8 + N;            // x in, y out; young; mask
  } else if (kernel == 12) {
    // y = K x on the matrix cores (FP64 MFMA element contraction)
    d2d(d_r.p, d_f.p, n3_, stream_);
    launch = [this]() { launch_apply_k64_mfma(L_, kc_, d_young, d_free, d_r, d_y, stream_); };
    F = 2.0 * 24 * 24 * E;
    B = 2.0 * 3 * N * 8 + E * 8 + N;            // x in, y out; young; mask
  } else if (kernel == 10) {
    // the multigrid solve's PCG stencil: p = z + beta p_old (no dinv: z is the
    // V-cycle's result), q = K p, p.q partials; beta = 0 from the scalars
    d2d(d_r.p, d_f.p, n3_, stream_);
    PcgScalars s0{};
    s0.rho = 1.0;
    s0.tol = 0.0;
    s0.maxit = 1 << 30;
    *h_pcg.p = s0;
    h2d(d_pcg.p, h_pcg.p, 1, stream_);
    launch = [this]() {
      launch_pcg_stencil64(L_, kc_, d_young, d_free, d_r, nullptr, d_p[0], d_p[1], d_q, d_pcg,
                           d_partials, d_counter, stream_, false);
    };
    F = 2.0 * 24 * 24 * E;
    B = 2.0 * 3 * N * 8 + 2.0 * 3 * N * 8 + E * 8 + N;  // z, p_old in; p, q out; young; mask
  } else if (kernel == 11) {
    // the multigrid solve's vector update: alpha = rho / p.q from stencil
    // partials (ones; rho = 0 keeps the vectors fixed), x, r updated (x as a
    // compensated sum, as the solve runs it by default), r.r partials,
    // z0 = omega dinv r for the V-cycle's first sweep
    hook_xlo.alloc(n3_);
    HIPCHK(hipMemsetAsync(hook_xlo.p, 0, hook_xlo.bytes(), stream_));
    PcgScalars s0{};
    s0.rho = 0.0;
    s0.tol = 0.0;
    s0.maxit = 1 << 30;
    *h_pcg.p = s0;
    h2d(d_pcg.p, h_pcg.p, 1, stream_);
    const unsigned gs = (unsigned)std::min<long long>(stencil_blocks(L_), 4096);
    std::vector<double> ones(gs, 1.0);
    h2d(d_partials.p, ones.data(), gs, stream_);
    launch = [this, gs, &hook_xlo]() {
      launch_fpcg_update_alpha(n3_, d_x, d_r, d_p[1], d_q, d_pcg, d_partials, gs, d_partials2,
                               d_dinv, 1.0, d_y.p, stream_, hook_xlo.p);
    };
    F = 10.0 * 3 * N;
    B = (6.0 + 4.0) * 3 * N * 8;                // x, xlo, p, r, q, dinv in; x, xlo, r, z0 out
  } else if (kernel == 9) {
    // one Helmholtz filter pass (density mode): element loads, the six 1-D
    // transforms (FP64 GEMMs) with the spectral scaling, element read-back
    if (!helm_) throw Error(kNotReady, "helm_setup has not been called");
    launch = [&]() { helm_->apply(d_dens, d_e[E_XF], 0); };
    F = 2.0 * 2.0 * N * (L_.NX + L_.NY + L_.NZ);
    B = 2.0 * 8 * E;                            // x in, x_tilde out
  } else {
    throw Error(kInvalidArgument, "unknown kernel id");
  }
  launch();  // warm
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  HIPCHK(hipEventRecord(a, stream_));
  for (int i = 0; i < reps; ++i) launch();
  HIPCHK(hipEventRecord(b, stream_));
  HIPCHK(hipEventSynchronize(b));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (avg_ms) *avg_ms = ms / reps;
  if (flops) *flops = F;
  if (bytes) *bytes = B;
  solved_ = false;
}

}  // namespace icarat_hip
