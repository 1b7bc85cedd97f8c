// context.cpp -- device state and host orchestration of one SIMP problem.
#include "context.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>

#include "mixed.hpp"
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
  HIPCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  d_ke.alloc(576);
  h2d(d_ke.p, prob_.ke0(), 576, stream_);
  for (DevBuf<double> *b : {&d_f, &d_b, &d_dinv, &d_x, &d_r, &d_q, &d_p[0], &d_p[1], &d_u, &d_nv,
                            &d_y, &d_dirlat}) {
    b->alloc(n3_);
    HIPCHK(hipMemsetAsync(b->p, 0, b->bytes(), stream_));
  }
  d_dens.alloc(L_.nelx);
  d_young.alloc(L_.nelx);
  for (auto &b : d_e) {
    b.alloc(L_.nelx);
    HIPCHK(hipMemsetAsync(b.p, 0, b.bytes(), stream_));
  }
  const long long stencil_blocks =
      (long long)((L_.NX + 31) / 32) * ((L_.NY + 7) / 8) * ((L_.NZ + 3) / 4);
  const long long filter_blocks = (long long)((L_.nx + 31) / 32) * ((L_.ny + 7) / 8) * ((L_.nz + 3) / 4);
  const long long nparts = std::max({stencil_blocks, filter_blocks, 4096LL}) * 4;
  d_partials.alloc(nparts);
  d_counter.alloc(4);
  HIPCHK(hipMemsetAsync(d_counter.p, 0, d_counter.bytes(), stream_));
  d_pcg.alloc(1);
  h_pcg.alloc(1);
  d_scal.alloc(16);
  h_scal.alloc(16);
  d_oc.alloc(1);
  h_oc.alloc(1);
  sync();
  dev_ready_ = true;
  if (prob_.dof_ready()) upload_dof();
}

Context::~Context() {
  if (!dev_ready_) return;
  for (int k = 0; k < 2; ++k) {
    if (pcg_exec_[k]) (void)hipGraphExecDestroy(pcg_exec_[k]);
    if (pcg_graph_[k]) (void)hipGraphDestroy(pcg_graph_[k]);
  }
  mixed_.reset();
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
  h2d(d_numeq[0].p, prob_.fext().data(), numeq, stream_);
  launch_numeq_to_lattice(L_, d_dof, numeq, d_numeq[0], nullptr, d_f, stream_);
  // constrained entries of the lattice load are 0 by construction
  sync();
  solved_ = false;
  fext_dirty_ = false;
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
  launch_diag_inv(L_, d_ke, d_young, d_free, d_dinv, nullptr, stream_);
  density_ready_ = true;
  solved_ = false;
  if (mixed_) mixed_->invalidate();
}

void Context::apply_k(const double *u, double *y, int precision) {
  require_density();
  const int numeq = prob_.numeq();
  h2d(d_numeq[0].p, u, numeq, stream_);
  launch_numeq_to_lattice(L_, d_dof, numeq, d_numeq[0], nullptr, d_y, stream_);
  if (precision == ICARAT_HIP_FP64) {
    launch_apply_k64(L_, d_ke, d_young, d_free, d_y, d_q, stream_);
  } else if (precision == ICARAT_HIP_FP32 || precision == ICARAT_HIP_BF16) {
    if (!mixed_) mixed_.reset(new MixedSolver(*this));
    mixed_->apply(d_y, d_q, precision);
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

void Context::build_pcg_graph() {
  if (pcg_exec_[0]) return;
  for (int par = 0; par < 2; ++par) {
    HIPCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
      for (int j = 0; j < kBatch; ++j) {
        double *p_old = d_p[(j + par) & 1], *p_new = d_p[(j + par + 1) & 1];
        launch_pcg_stencil64(L_, d_ke, d_young, d_free, d_r, d_dinv, p_old, p_new, d_q, d_pcg,
                             d_partials, d_counter, stream_);
        launch_pcg_update64(n3_, d_x, d_r, p_new, d_q, d_dinv, d_pcg, d_partials, d_counter,
                            stream_);
      }
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(stream_, &pcg_graph_[par]));
    HIPCHK(hipGraphInstantiate(&pcg_exec_[par], pcg_graph_[par], nullptr, nullptr, 0));
  }
}

PcgScalars Context::pcg64(double tol, int maxit, bool warm) {
  build_pcg_graph();
  PcgScalars init{};
  init.tol = tol;
  init.maxit = maxit;
  init.delta = 0.0;
  *h_pcg.p = init;
  h2d(d_pcg.p, h_pcg.p, 1, stream_);
  if (warm) d2d(d_nv.p, d_x.p, n3_, stream_);
  // x = 0, r = b, p = 0, bnorm2 = b.b
  launch_pcg_init64(n3_, d_b, d_x, d_r, d_p[0], d_dinv, d_pcg, d_partials, d_counter, stream_);
  if (warm) {
    d2d(d_x.p, d_nv.p, n3_, stream_);
    launch_apply_k64(L_, d_ke, d_young, d_free, d_x, d_y, stream_);
    launch_replace64(n3_, d_b, d_y, d_r, d_dinv, d_pcg, d_partials, stream_);
  }
  d2h(h_pcg.p, d_pcg.p, 1, stream_);
  sync();
  replacements_ = 0;
  while (!h_pcg.p->done || h_pcg.p->status == 0) {
    // iterate until the recursive stop test, a cap or a breakdown
    const long long max_batches = (long long)(maxit - h_pcg.p->iter) / kBatch + 3;
    for (long long b = 0; b < max_batches && !h_pcg.p->done; ++b) {
      HIPCHK(hipGraphLaunch(pcg_exec_[h_pcg.p->iter & 1], stream_));
      d2h(h_pcg.p, d_pcg.p, 1, stream_);
      sync();
    }
    if (!h_pcg.p->done) throw Error(kInternal, "PCG did not reach its stop test");
    if (h_pcg.p->status != 0) break;
    // verified stop: the FP64 true residual must meet the tolerance too;
    // otherwise it replaces the recursive one and the iteration continues
    launch_apply_k64(L_, d_ke, d_young, d_free, d_x, d_y, stream_);
    launch_replace64(n3_, d_b, d_y, d_r, d_dinv, d_pcg, d_partials, stream_);
    d2h(h_pcg.p, d_pcg.p, 1, stream_);
    sync();
    if (h_pcg.p->done) break;
    if (++replacements_ > 20 || h_pcg.p->iter > maxit) {
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
  launch_apply_k64(L_, d_ke, d_young, d_free, d_u, d_y, stream_);
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
  // right-hand side: fext minus the load of the prescribed displacements
  d2d(d_b.p, d_f.p, n3_, stream_);
  if (prob_.has_nonzero_dirichlet()) {
    launch_apply_k64(L_, d_ke, d_young, d_free, d_dirlat, d_y, stream_);
    launch_axpby(n3_, -1.0, d_y, 1.0, d_b, stream_);
  }
  icarat_solve_stats S{};
  S.status = ICARAT_HIP_OK;
  if (o.precision == ICARAT_HIP_FP64 && o.precond == ICARAT_HIP_PRECOND_JACOBI) {
    PcgScalars r = pcg64(o.tol, o.maxit);
    S.iters = r.iter;
    S.refine_steps = replacements_;
    S.rec_relres = r.bnorm2 > 0 ? std::sqrt(r.rr) / std::sqrt(r.bnorm2) : 0.0;
    if (r.status == 2) throw Error(kNumerical, "PCG breakdown (non-positive curvature or NaN)");
    if (r.status == 1) S.status = ICARAT_HIP_NOT_CONVERGED;
  } else {
    if (!mixed_) mixed_.reset(new MixedSolver(*this));
    mixed_->solve(o, S);
  }
  assemble_nodal_solution();
  S.true_relres = true_residual_device();
  S.seconds = now_s() - t0;
  if (u_host) {
    const int numeq = prob_.numeq();
    launch_lattice_to_numeq(L_, d_dof, numeq, d_u, d_numeq[0], stream_);
    d2h(u_host, d_numeq[0].p, numeq, stream_);
    sync();
  }
  solved_ = true;
  if (st) *st = S;
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
void Context::simp_run(const icarat_simp_opts &o, double *x, double *hist, icarat_simp_stats *st) {
  require_dof();
  ensure_fext();
  if (!filter_ready_) throw Error(kNotReady, "filter_setup has not been called");
  if (o.n_iter < 0) throw Error(kInvalidArgument, "n_iter must be >= 0");
  if (o.optimizer != 0 && o.optimizer != 1) throw Error(kInvalidArgument, "optimizer must be 0 (OC) or 1 (MMA)");
  const auto &g = prob_.spec();
  const long long n = L_.nelx;
  icarat_simp_stats S{};
  sync();
  const double t_all = now_s();
  h2d(d_e[E_DES].p, x, n, stream_);
  if (o.optimizer == 1) {
    if (!mma_) mma_.reset(new DeviceMMA(*this));
    mma_->reset();
  }
  double V0 = 0.0;
  for (int step = 1; step <= o.n_iter; ++step) {
    double t0 = now_s();
    // x_tilde = filter(x); x_hat = projection(x_tilde) or x_tilde
    filter_device(d_e[E_DES], d_e[E_XF], 0);
    const double beta = std::min(std::pow(2, (int)(step / 40)), 16.0);
    double *xp = d_e[E_XF];
    if (o.use_projection) {
      launch_threshold(n, d_e[E_XF], beta, 0.5, d_e[E_XP], 0, stream_);
      launch_threshold(n, d_e[E_XF], beta, 0.5, d_e[E_THD], 1, stream_);
      xp = d_e[E_XP];
    }
    d2d(d_dens.p, xp, n, stream_);
    set_density_device();
    sync();
    double t1 = now_s();
    S.filter_s += t1 - t0;
    icarat_solve_stats ss{};
    solve(o.solve, nullptr, &ss);
    S.pcg_iters += ss.iters;
    S.max_true_relres = std::max(S.max_true_relres, ss.true_relres);
    double t2 = now_s();
    S.solve_s += t2 - t1;
    // compliance and volume
    launch_dot(n3_, d_f, d_u, d_scal, d_partials, d_counter, stream_);
    launch_volume(n, d_dens, vol_e_, d_scal.p + 1, d_partials, d_counter, stream_);
    d2h(h_scal.p, d_scal.p, 2, stream_);
    sync();
    const double c = h_scal.p[0], vol = h_scal.p[1];
    if (step == 1) V0 = vol;
    launch_flush_small(n3_, d_u, d_nv, stream_);
    launch_sensitivity64(L_, d_ke, d_dens, d_nv, g.young1, g.young2, g.penalty, vol_e_,
                         o.optimizer == 1 ? V0 : 1.0, d_e[E_DFDS], d_e[E_DGDS], stream_);
    if (o.use_projection) {
      launch_mul(n, d_e[E_THD], d_e[E_DFDS], stream_);
      launch_mul(n, d_e[E_THD], d_e[E_DGDS], stream_);
    }
    sync();
    double t3 = now_s();
    S.sens_s += t3 - t2;
    filter_device(d_e[E_DFDS], d_e[E_TMP], 1);
    std::swap(d_e[E_DFDS], d_e[E_TMP]);
    filter_device(d_e[E_DGDS], d_e[E_TMP], 1);
    std::swap(d_e[E_DGDS], d_e[E_TMP]);
    sync();
    double t4 = now_s();
    S.filter_s += t4 - t3;
    if (hist) {
      hist[4 * (step - 1) + 0] = c;
      hist[4 * (step - 1) + 1] = vol;
      hist[4 * (step - 1) + 2] = ss.iters;
      hist[4 * (step - 1) + 3] = ss.true_relres;
    }
    if (o.optimizer == 1) {
      const double gc = step == 1 ? 0.0 : vol / V0 - 1.0;
      mma_->update(d_e[E_DES], d_e[E_DFDS], gc, d_e[E_DGDS], o.xmin, o.xmax);
    } else {
      oc_device(d_e[E_DES], d_e[E_DFDS], d_e[E_DGDS], V0, o.xmin, o.xmax, o.use_projection,
                beta);
    }
    sync();
    S.update_s += now_s() - t4;
  }
  d2h(x, d_e[E_DES].p, n, stream_);
  sync();
  S.seconds = now_s() - t_all;
  if (st) *st = S;
}

// ---- measurement hook ------------------------------------------------------
void Context::time_kernel(int kernel, int reps, double *avg_ms, double *flops, double *bytes) {
  require_density();
  ensure_fext();
  if (reps <= 0) throw Error(kInvalidArgument, "reps must be positive");
  const double N = (double)L_.numnp, E = (double)L_.nelx;
  double F = 0.0, B = 0.0;
  std::function<void()> launch;
  if (kernel == 0) {
    // p = dinv r + beta p_old with r = f (non-zero) and beta = 0, so p.q > 0
    // and the stop flag never rises between launches
    d2d(d_r.p, d_f.p, n3_, stream_);
    PcgScalars s0{};
    s0.rho = 1.0;
    s0.tol = 0.0;
    s0.maxit = 1 << 30;
    *h_pcg.p = s0;
    h2d(d_pcg.p, h_pcg.p, 1, stream_);
    launch = [&]() {
      launch_pcg_stencil64(L_, d_ke, d_young, d_free, d_r, d_dinv, d_p[0], d_p[1], d_q, d_pcg,
                           d_partials, d_counter, stream_, false);
    };
    F = 2.0 * 24 * 24 * E;                      // dense KE0 x U_e contraction
    B = 3.0 * N * 8 * 3 + 2.0 * 3 * N * 8 + E * 8 + N;  // r, dinv, p_old in; p, q out; young; mask
  } else if (kernel == 1) {
    PcgScalars s0{};
    s0.rho = 1.0;
    s0.alpha = 0.0;
    s0.tol = 0.0;
    s0.maxit = 1 << 30;
    *h_pcg.p = s0;
    h2d(d_pcg.p, h_pcg.p, 1, stream_);
    launch = [&]() {
      launch_pcg_update64(n3_, d_x, d_r, d_p[1], d_q, d_dinv, d_pcg, d_partials, d_counter,
                          stream_, false);
    };
    F = 8.0 * 3 * N;
    B = (5.0 + 2.0) * 3 * N * 8;                // x, r, p, q, dinv in; x, r out
  } else if (kernel == 2) {
    if (!filter_ready_) throw Error(kNotReady, "filter_setup has not been called");
    launch = [&]() { launch_density_filter(L_, ftab_, d_dens, d_e[2], d_e[3], 0, stream_); };
    F = 4.0 * ftab_.count * E;
    B = 3.0 * 8 * E;                            // x in; x_tilde, aa out
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
