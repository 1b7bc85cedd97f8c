// multigrid_cases.cpp -- the load cases of one problem in one multigrid-PCG
// solve (see multigrid.hpp).  The reference solves each load case of its
// multi-load compliance on its own and sums f_c . u_c
// (sample/software/opt-multi-load/opt_multi_load.cpp:118-146); here the cases
// share every launch of a batch: K [p_1 .. p_k] in one z-marching pass, one
// update over the stacked vectors, one V- or W-cycle with every transfer,
// sweep and dense solve taking the cases together.  Each case keeps its own
// scalars, partials (in the slots of a single-case launch, so its sums are
// those of its own solve), verified stop and endgame.
#include <algorithm>
#include <cmath>

#include "context.hpp"
#include "learned_kernels.hpp"
#include "mg_kernels.hpp"
#include "multigrid.hpp"

namespace icarat_hip {

namespace {
constexpr int kRangeKey = 256;  // graph key: first case x kRangeKey + case count
constexpr long long kVecParts = 4096;  // room for the vector kernels' partials per quantity
}  // namespace

bool Multigrid::cases_capable(const icarat_solve_opts &o) const {
  if (o.precision != ICARAT_HIP_FP64 || galerkin_ || overlap_) return false;
  if ((batch_iters() & 1) != 0) return false;  // p buffers alternate at fixed parity
  if (wlev_ > 0 && o_.w_switch != 0.0) return false;  // the cycle shape switches per case
  if (f_.kc_.kernel != 0) return false;  // the multi-case PCG stencil is z-marching
  for (int l = 0; l < levels(); ++l) {
    const bool dense_level = l == levels() - 1 && dense_;
    if (!dense_level && !direct_[l] && const_cast<Multigrid *>(this)->level(l).kc_.kernel != 0)
      return false;
  }
  return true;
}

void Multigrid::mc_drop_graphs() {
  for (auto &e : mc_.exec)
    if (e.second) (void)hipGraphExecDestroy(e.second);
  for (auto &g : mc_.graph)
    if (g.second) (void)hipGraphDestroy(g.second);
  mc_.exec.clear();
  mc_.graph.clear();
}

void Multigrid::mc_alloc(int nc) {
  Cases &m = mc_;
  if (m.nc == nc) return;
  mc_drop_graphs();
  hipStream_t s = f_.stream();
  const int L = levels();
  m.b.resize(L);
  m.x.resize(L);
  m.y.resize(L);
  m.r.resize(L);
  m.u.resize(L);
  m.nv.resize(L);
  std::vector<DevBuf<double> *> all;
  for (int l = 0; l < L; ++l) {
    const size_t n = (size_t)nc * level(l).n3_;
    for (auto *v : {&m.x, &m.y, &m.r}) {
      (*v)[l].alloc(n);
      all.push_back(&(*v)[l]);
    }
    if (l > 0)
      for (auto *v : {&m.b, &m.u, &m.nv}) {
        (*v)[l].alloc(n);
        all.push_back(&(*v)[l]);
      }
  }
  const size_t n0 = (size_t)nc * f_.n3_;
  for (DevBuf<double> *v : {&m.R, &m.P[0], &m.P[1], &m.Q, &m.yt}) {
    v->alloc(n0);
    all.push_back(v);
  }
  if (o_.compensated) {
    m.xlo.alloc(n0);
    all.push_back(&m.xlo);
  }
  m.xbest.release();
  m.S.alloc(nc);
  m.hS.alloc(nc);
  m.pst.alloc((size_t)nc * stencil_blocks(f_.lattice()));
  m.pup.alloc((size_t)nc * kVecParts);
  m.pz.alloc((size_t)nc * 2 * std::max(stencil_blocks(f_.lattice()), kVecParts));
  if (dense_) m.dpart.alloc((size_t)nc * mg_dense_part_size(level(L - 1).lattice()));
  // the lattice padding (Np rounds the node count up) is read by the vector
  // kernels and never written: it holds zeros, as in the single-case buffers
  for (DevBuf<double> *v : all) HIPCHK(hipMemsetAsync(v->p, 0, v->bytes(), s));
  f_.sync();
  m.nc = nc;
  m.zres = nullptr;
}

double *Multigrid::cycle_mc(int l, const double *b, bool x0_ready, ZDot *zd, int c0, int ncs) {
  Cases &m = mc_;
  hipStream_t s = f_.stream();
  Context &c = level(l);
  const Lattice &L = c.lattice();
  const long long n = c.n3_, o = (long long)c0 * n;
  double *x = m.x[l].p, *y = m.y[l].p, *r = m.r[l].p;
  const double w = omega_[l];
  if (l == levels() - 1 && dense_) {
    launch_mg_dense_solve(L, ainv_, b + o, x + o, m.dpart, s, ncs, n);
    return x;
  }
  double *cur = x, *oth = y;
  auto sweep = [&]() {
    if (direct_[l])
      launch_mg_direct(MG_JACOBI, L, c.kc_, c.d_young, c.d_free, cur + o, b + o, c.d_dinv, w,
                       oth + o, s, ncs, n);
    else
      launch_apply_k64_mc(L, c.kc_, c.d_young, c.d_free, cur + o, oth + o, b + o, c.d_dinv, w,
                          nullptr, nullptr, ncs, n, s);
    std::swap(cur, oth);
  };
  if (l == 0 && !x0_ready && starts_swept(0))
    launch_mg_jacobi(n, w, c.d_dinv, b + o, nullptr, x + o, s, ncs);
  if (l == levels() - 1) {
    for (int i = 1; i < std::max(1, o_.nu_coarse); ++i) sweep();
    return cur;
  }
  const int pre = std::max(0, o_.nu_pre);
  for (int i = 1; i < pre; ++i) sweep();
  Context &cc = level(l + 1);
  const Lattice &Lc = cc.lattice();
  const long long nc = cc.n3_, oc = (long long)c0 * nc;
  double *bc = m.b[l + 1].p, *xc0 = m.x[l + 1].p;
  {
    const double *dinv_c = starts_swept(l + 1) ? cc.d_dinv.p : nullptr;
    const double *rf = b + o, *yf = nullptr;
    if (pre > 0 && direct_[l]) {
      launch_mg_direct(MG_RESIDUAL, L, c.kc_, c.d_young, c.d_free, cur + o, b + o, nullptr, 0.0,
                       r + o, s, ncs, n);
      rf = r + o;
    } else if (pre > 0) {
      launch_apply_k64_mc(L, c.kc_, c.d_young, c.d_free, cur + o, oth + o, nullptr, nullptr, 0.0,
                          nullptr, nullptr, ncs, n, s);
      yf = oth + o;
    }
    launch_mg_restrict(L, Lc, rf, yf, cc.d_free, bc + oc, dinv_c, omega_[l + 1], xc0 + oc, s, ncs,
                       n, nc);
  }
  const double *xc = cycle_mc(l + 1, bc, false, nullptr, c0, ncs);
  const double *xc2 = nullptr;
  if (l + 1 <= wcur_ && l + 1 < levels() - 1) {
    // the W-cycle's second coarse visit on the remaining coarse residual
    double *u = m.u[l + 1].p, *nv = m.nv[l + 1].p;
    double *x0 = starts_swept(l + 1) ? xc0 + oc : nullptr;
    if (direct_[l + 1] && xc != xc0) {
      launch_mg_direct_wres(Lc, cc.kc_, cc.d_young, cc.d_free, xc + oc, bc + oc, cc.d_dinv,
                            omega_[l + 1], u + oc, nv + oc, x0, s, ncs, nc);
    } else if (!direct_[l + 1] && xc != xc0) {
      launch_apply_k64_mc(Lc, cc.kc_, cc.d_young, cc.d_free, xc + oc, nv + oc, bc + oc, cc.d_dinv,
                          omega_[l + 1], nullptr, nullptr, ncs, nc, s, u + oc, x0);
    } else {
      launch_copy64(ncs * nc, xc + oc, u + oc, s);
      if (direct_[l + 1]) {
        launch_mg_direct(MG_RESIDUAL, Lc, cc.kc_, cc.d_young, cc.d_free, xc + oc, bc + oc, nullptr,
                         0.0, nv + oc, s, ncs, nc);
      } else {
        launch_apply_k64_mc(Lc, cc.kc_, cc.d_young, cc.d_free, xc + oc, nv + oc, nullptr, nullptr,
                            0.0, nullptr, nullptr, ncs, nc, s);
        launch_axpby(ncs * nc, 1.0, bc + oc, -1.0, nv + oc, s);
      }
      if (x0) launch_mg_jacobi(nc, omega_[l + 1], cc.d_dinv, nv + oc, nullptr, x0, s, ncs);
    }
    xc2 = cycle_mc(l + 1, nv, false, nullptr, c0, ncs);
    xc = u;
  }
  const int post = std::max(0, o_.nu_post);
  const bool fused = fuse_ && direct_[l] && pre > 0 && post > 0;
  int first = 0;
  if (fused) {
    launch_mg_prolong_jacobi(L, Lc, c.kc_, c.d_young, c.d_free, xc + oc, cur + o, b + o, c.d_dinv,
                             w, oth + o, s, xc2 ? xc2 + oc : nullptr, ncs, n, nc);
    std::swap(cur, oth);
    first = 1;
  } else {
    launch_mg_prolong_add(L, Lc, xc + oc, c.d_free, cur + o, s, pre > 0, xc2 ? xc2 + oc : nullptr,
                          ncs, n, nc);
  }
  for (int i = first; i < post; ++i) {
    if (i == post - 1 && l == 0 && !direct_[l]) {
      // the last fine sweep out of place, with the z.r and z.q partials
      const double *q = zd ? zd->q : m.Q.p;
      double *part = zd ? zd->part : m.pz.p;
      const unsigned cnt = launch_apply_k64_mc(L, c.kc_, c.d_young, c.d_free, cur + o, oth + o,
                                               b + o, c.d_dinv, w, q + o, part, ncs, n, s);
      if (zd) zd->count = cnt;
      std::swap(cur, oth);
    } else {
      sweep();
    }
  }
  return cur;
}

void Multigrid::mc_enqueue_batch(int lo, int cnt) {
  Cases &m = mc_;
  Context &c = f_;
  hipStream_t s = c.stream();
  const Lattice &L = c.lattice();
  const long long n = c.n3_, o = (long long)lo * n;
  const bool x0_fused = starts_swept(0);
  for (int j = 0; j < batch_iters(); ++j) {
    double *p_old = m.P[j & 1].p + o, *p_new = m.P[(j + 1) & 1].p + o;
    // p = z + beta p_old, q = K p, p.q partials
    const unsigned gs = launch_pcg_stencil64_mcz(L, c.kc_, c.d_young, c.d_free, m.zres + o, p_old,
                                                 p_new, m.Q.p + o, m.S.p + lo, m.pst, cnt, n, s);
    // alpha, x and r, r.r partials, z0 = omega D^-1 r
    const unsigned gu = launch_fpcg_update_alpha(
        n, m.X + o, m.R.p + o, p_new, m.Q.p + o, m.S.p + lo, m.pst, gs, m.pup, c.d_dinv, omega_[0],
        x0_fused ? m.x[0].p + o : nullptr, s, m.xlo.p ? m.xlo.p + o : nullptr, cnt, n);
    ZDot zd{m.Q.p, m.pz.p, 0};
    const double *z = cycle_mc(0, m.R.p, x0_fused, &zd, lo, cnt);
    if (z != m.zres) throw Error(kInternal, "multigrid cases: unexpected cycle result buffer");
    if (!zd.count)
      zd.count = launch_zdot_partials(n, m.zres + o, m.R.p + o, m.Q.p + o, m.pz, s, cnt, n);
    if ((size_t)gs * cnt > m.pst.n || (size_t)gu * cnt > m.pup.n ||
        (size_t)2 * zd.count * cnt > m.pz.n)
      throw Error(kInternal, "multigrid cases: partials buffer too small");
    // stop tests, then rho and beta
    launch_fpcg_check_beta(m.pup, gu, m.pz, zd.count, m.S.p + lo, s, cnt);
  }
}

void Multigrid::mc_restart(int lo, int cnt) {
  Cases &m = mc_;
  Context &c = f_;
  hipStream_t s = c.stream();
  const long long n = c.n3_;
  for (int k = lo; k < lo + cnt; ++k) launch_fpcg_restart_flag(m.S.p + k, s);
  double *z = cycle_mc(0, m.R.p, false, nullptr, lo, cnt);
  if (m.zres && z != m.zres) throw Error(kInternal, "multigrid cases: unexpected cycle result buffer");
  m.zres = z;
  for (int k = lo; k < lo + cnt; ++k)
    launch_zdot_beta(n, z + k * n, m.R.p + k * n, m.Q.p + k * n, m.S.p + k, c.d_partials, s);
}

void Multigrid::mc_begin(const icarat_solve_opts &o, int nc, const double *B, double *X,
                         const char *guess) {
  if (!cases_capable(o)) throw Error(kInternal, "multigrid cases: unsupported configuration");
  if (nc < 1 || nc >= kRangeKey) throw Error(kInvalidArgument, "load case count out of range");
  Context &c = f_;
  hipStream_t s = c.stream();
  set_fine_fp32(false);
  refresh_now();
  mc_alloc(nc);
  Cases &m = mc_;
  if (m.X != X) {
    mc_drop_graphs();
    m.X = X;
  }
  m.B = B;
  m.wsave = wcur_;
  wcur_ = wlev_;  // the first phase's cycle for the whole solve (w_switch 0)
  tol_ = o.tol;
  const long long n = c.n3_;
  const Lattice &L = c.lattice();
  // per case: x = 0 (or the guess), r = b (or b - K x), scalars
  for (int k = 0; k < nc; ++k) {
    PcgScalars s0{};
    s0.tol = o.tol;
    s0.maxit = o.maxit;
    m.hS.p[k] = s0;
  }
  HIPCHK(hipMemcpyAsync(m.S.p, m.hS.p, sizeof(PcgScalars) * nc, hipMemcpyHostToDevice, s));
  if (m.xlo.p) HIPCHK(hipMemsetAsync(m.xlo.p, 0, m.xlo.bytes(), s));
  for (int k = 0; k < nc; ++k) {
    double *xk = X + k * n, *rk = m.R.p + k * n, *yk = m.yt.p + k * n;
    const double *bk = B + k * n;
    const bool g = guess && guess[k];
    if (g) launch_copy64(n, xk, yk, s);
    launch_pcg_init64(n, bk, xk, rk, m.P[0].p + k * n, c.d_dinv, m.S.p + k, c.d_partials,
                      c.d_counter, s);
    if (g) {
      launch_copy64(n, yk, xk, s);
      launch_apply_k64(L, c.kc_, c.d_young, c.d_free, xk, yk, s);
      launch_replace64(n, bk, yk, rk, c.d_dinv, m.S.p + k, c.d_partials, s);
    }
  }
  mc_restart(0, nc);
  HIPCHK(hipMemcpyAsync(m.hS.p, m.S.p, sizeof(PcgScalars) * nc, hipMemcpyDeviceToHost, s));
  c.sync();
  // 0 running, 1 finished, 2 left to the caller
  m.state.assign(nc, 0);
  m.repl.assign(nc, 0);
  m.since.assign(nc, 0);
  m.verifs.assign(nc, 0);
  m.best.assign(nc, INFINITY);
  m.redo.assign(nc, 0);
  m.max_batches = ((long long)o.maxit / batch_iters() + 3) * (kMaxVerifications + 2) + 16;
  m.batches = 0;
  m.lo = 0;
  m.cnt = 0;
}

bool Multigrid::mc_poll() {
  Cases &m = mc_;
  Context &c = f_;
  hipStream_t s = c.stream();
  const int nc = m.nc;
  const long long n = c.n3_;
  const Lattice &L = c.lattice();
  auto put = [&](int k) {
    HIPCHK(hipMemcpyAsync(m.S.p + k, m.hS.p + k, sizeof(PcgScalars), hipMemcpyHostToDevice, s));
  };
  for (;;) {
    // cases whose batch ended: the FP64 true residual decides (verified stop)
    std::vector<int> ver;
    for (int k = 0; k < nc; ++k) {
      if (m.state[k] != 0) continue;
      PcgScalars &h = m.hS.p[k];
      if (!h.done) continue;
      if (h.status == 4 && wcur_ > 0) {
        // the W-cycle broke down: its own solve goes on with the V-cycle
        m.state[k] = 2;
        m.redo[k] = 1;
        continue;
      }
      if (h.status != 0) {
        m.state[k] = 1;  // iteration cap or breakdown
        continue;
      }
      double *xk = m.X + k * n, *yk = m.yt.p + k * n;
      launch_apply_k64(L, c.kc_, c.d_young, c.d_free, xk, yk, s);
      launch_replace64(n, m.B + k * n, yk, m.R.p + k * n, c.d_dinv, m.S.p + k, c.d_partials, s);
      mc_restart(k, 1);
      HIPCHK(hipMemcpyAsync(m.hS.p + k, m.S.p + k, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
      ver.push_back(k);
    }
    if (!ver.empty()) c.sync();
    for (int k : ver) {
      ++m.repl[k];
      PcgScalars &h = m.hS.p[k];
      if (h.done) {
        m.state[k] = 1;  // the verified residual passes the recursive target
        continue;
      }
      const double bnorm = std::sqrt(h.bnorm2);
      const double tr = bnorm > 0.0 ? std::sqrt(h.rr) / bnorm : 0.0;
      if (tr < tol_) {
        // passes the solve's tolerance; the batches skip a done case
        h.done = 1;
        h.status = 0;
        put(k);
        m.state[k] = 1;
        continue;
      }
      ++m.verifs[k];
      if (tr < m.best[k]) {
        m.best[k] = tr;
        m.since[k] = 0;
        if (!m.xbest.p) m.xbest.alloc((size_t)nc * n);
        launch_copy64(n, m.X + k * n, m.xbest.p + k * n, s);
      } else {
        ++m.since[k];
      }
      if (m.since[k] < kStallVerifications && m.verifs[k] <= kMaxVerifications) {
        // the endgame target of after_verify()
        const double t0 = h.tol;
        const double g2 = std::max(0.0, tr * tr - t0 * t0);
        const double t =
            std::min(t0, std::sqrt(std::max(0.81 * tol_ * tol_ - g2, 0.09 * tol_ * tol_)));
        if (t < t0) {
          h.tol = t;
          put(k);
        }
        continue;
      }
      // stalled at the floor: the best verified iterate, not converged
      launch_copy64(n, m.xbest.p + k * n, m.X + k * n, s);
      if (m.xlo.p) HIPCHK(hipMemsetAsync(m.xlo.p + k * n, 0, sizeof(double) * n, s));
      h.done = 1;
      h.status = 1;
      put(k);
      m.state[k] = 1;
    }
    int lo = nc, hi = -1;
    for (int k = 0; k < nc; ++k)
      if (m.state[k] == 0) {
        lo = std::min(lo, k);
        hi = std::max(hi, k);
      }
    if (hi < 0) return true;
    // a case that has just verified may be done already: handle it first
    bool pending = false;
    for (int k = lo; k <= hi; ++k)
      if (m.state[k] == 0 && m.hS.p[k].done) pending = true;
    if (pending) continue;
    // the next batch runs over the range of running cases (cases in it that
    // are done are skipped by the PCG kernels; their cycle work is discarded)
    m.lo = lo;
    m.cnt = hi - lo + 1;
    return false;
  }
}

void Multigrid::mc_build_graph() {
  Cases &m = mc_;
  Context &c = f_;
  if (!c.use_graphs_ || m.cnt < 1) return;
  const int key = m.lo * kRangeKey + m.cnt;
  if (m.exec.count(key)) return;
  hipStream_t s = c.stream();
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    mc_enqueue_batch(m.lo, m.cnt);
  } catch (...) {
    hipGraph_t gg = nullptr;
    (void)hipStreamEndCapture(s, &gg);
    if (gg) (void)hipGraphDestroy(gg);
    throw;
  }
  HIPCHK(hipStreamEndCapture(s, &g));
  hipGraphExec_t e = nullptr;
  HIPCHK(hipGraphInstantiate(&e, g, nullptr, nullptr, 0));
  m.graph[key] = g;
  m.exec[key] = e;
}

void Multigrid::mc_launch() {
  Cases &m = mc_;
  Context &c = f_;
  hipStream_t s = c.stream();
  if (m.cnt < 1) throw Error(kInternal, "multigrid cases: no running case to launch");
  if (++m.batches > m.max_batches) throw Error(kInternal, "multigrid PCG did not reach its stop test");
  if (c.use_graphs_) {
    mc_build_graph();
    HIPCHK(hipGraphLaunch(m.exec.at(m.lo * kRangeKey + m.cnt), s));
  } else {
    mc_enqueue_batch(m.lo, m.cnt);
  }
  HIPCHK(hipMemcpyAsync(m.hS.p + m.lo, m.S.p + m.lo, sizeof(PcgScalars) * m.cnt,
                        hipMemcpyDeviceToHost, s));
}

void Multigrid::mc_end(icarat_solve_stats *st, std::vector<char> &redo) {
  Cases &m = mc_;
  wcur_ = m.wsave;
  finish();
  redo = m.redo;
  for (int k = 0; k < m.nc; ++k) {
    if (m.redo[k]) continue;
    const PcgScalars &h = m.hS.p[k];
    if (h.done && (h.status == 2 || h.status == 4))
      throw Error(kNumerical, "multigrid PCG breakdown (the V-cycle is not positive definite here)");
    icarat_solve_stats S{};
    S.iters = h.iter;
    S.refine_steps = m.repl[k];
    const double bnorm = std::sqrt(h.bnorm2);
    S.rec_relres = bnorm > 0.0 ? std::sqrt(h.rr) / bnorm : 0.0;
    S.status = (h.done && h.status == 0) || bnorm == 0.0 ? ICARAT_HIP_OK : ICARAT_HIP_NOT_CONVERGED;
    if (st) st[k] = S;
  }
}

void Multigrid::solve_cases(const icarat_solve_opts &o, int nc, const double *B, double *X,
                            const char *guess, icarat_solve_stats *st, std::vector<char> &redo) {
  mc_begin(o, nc, B, X, guess);
  try {
    while (!mc_poll()) {
      mc_launch();
      f_.sync();
    }
  } catch (...) {
    wcur_ = mc_.wsave;
    throw;
  }
  mc_end(st, redo);
}

}  // namespace icarat_hip
