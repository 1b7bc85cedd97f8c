// multigrid.cpp -- see multigrid.hpp.
#include "multigrid.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>

#include "context.hpp"
#include "galerkin_kernels.hpp"
#include "learned_kernels.hpp"
#include "mg_kernels.hpp"

namespace icarat_hip {

namespace {
// mu = lambda_max(diag(KE0)^-1 KE0) by power iteration on the similar symmetric
// matrix diag^-1/2 KE0 diag^-1/2.  Element by element, x^T K x = sum_e E_e
// x_e^T KE0 x_e <= mu sum_e E_e x_e^T diag(KE0) x_e = mu x^T D x, so mu bounds
// lambda_max(D^-1 K) for every density field, and Jacobi damped by omega with
// omega mu < 2 is a convergent (energy-norm contracting) smoother.
double jacobi_bound(const double *ke) {
  double s[24], v[24], w[24];
  for (int i = 0; i < 24; ++i) {
    s[i] = 1.0 / std::sqrt(ke[i * 24 + i]);
    v[i] = 1.0 + 0.01 * i;
  }
  double lam = 0.0;
  for (int it = 0; it < 500; ++it) {
    double nw = 0.0;
    for (int i = 0; i < 24; ++i) {
      double a = 0.0;
      for (int j = 0; j < 24; ++j) a += s[i] * ke[i * 24 + j] * s[j] * v[j];
      w[i] = a;
      nw += a * a;
    }
    nw = std::sqrt(nw);
    double nv = 0.0;
    for (int i = 0; i < 24; ++i) nv += v[i] * v[i];
    lam = nw / std::sqrt(nv);
    for (int i = 0; i < 24; ++i) v[i] = w[i] / nw;
  }
  return lam;
}
}  // namespace

Multigrid::Multigrid(Context &fine, const icarat_mg_opts &o) : f_(fine), o_(o) {
  f_.ensure_device();
  if (const char *e = std::getenv("ICARAT_HIP_MG_FUSE")) fuse_ = std::string(e) != "0";
  if (const char *e = std::getenv("ICARAT_HIP_MG_OVERLAP")) overlap_ = std::string(e) == "1";
  const GridSpec g0 = f_.problem().spec();
  GridSpec g = g0;
  auto dofs = [](const GridSpec &q) { return 3LL * (q.nx + 1) * (q.ny + 1) * (q.nz + 1); };
  // automatic (dense_dofs < 0): up to 1/3000 of the fine lattice DOFs, at
  // least 1024.  Measured on the SIMP loops: at 256x128x128 (12.8M DOFs) the
  // 16x8x8 grid (4131 DOFs) solved exactly instead of coarsened to 8x4x4
  // halves the PCG iterations and the time per design iteration; at 128x64x64
  // (1.6M DOFs) the same dense level costs 1.5x the time of the 8x4x4 one
  // (its 136 MB inverse is read at every visit).
  const long long dense_max = std::min<long long>(
      o_.dense_dofs >= 0 ? o_.dense_dofs : std::max(1024LL, dofs(g0) / 3000), MG_DENSE_MAX);
  // coarsen while every dimension halves exactly and the grid keeps some size,
  // up to the first grid small enough for the exact solve
  auto halves = [](const GridSpec &q) {
    return q.nx % 2 == 0 && q.ny % 2 == 0 && q.nz % 2 == 0 &&
           (long long)(q.nx / 2) * (q.ny / 2) * (q.nz / 2) >= 8;
  };
  while ((int)lv_.size() + 1 < std::max(1, o_.max_levels) && halves(g) &&
         !(lv_.size() > 0 && dofs(g) <= dense_max)) {
    g.nx /= 2;
    g.ny /= 2;
    g.nz /= 2;
    std::unique_ptr<Context> c(new Context(g));
    for (const auto &b : f_.problem().dirichlet_boxes()) c->problem().add_dirichlet(b.flag, b.val, b.range);
    c->make_dof();
    c->ensure_device();
    lv_.push_back(std::move(c));
  }
  galerkin_ = o_.coarse_avg == 5 && levels() > 1;
  for (int l = 0; l < levels(); ++l)
    omega_.push_back(galerkin_ && l > 0 ? 1.0  // folded into the level's dinv
                                        : o_.omega / jacobi_bound(level(l).problem().ke0()));
  // smoothing kernels per level: the fused one-thread-per-node sweep where the
  // z-marching launch is latency-bound (every Galerkin level: their operators
  // are explicit element matrices); the smallest coarsest grids exactly
  for (int l = 0; l < levels(); ++l)
    direct_.push_back((galerkin_ && l > 0) ||
                              level(l).lattice().numnp <= (long long)std::max(0, o_.direct_nodes)
                          ? 1
                          : 0);
  if (galerkin_) {
    kmat_.resize(levels());
    long long maxe = 0;
    for (int l = 1; l < levels(); ++l) {
      kmat_[l].alloc(576 * (size_t)level(l).lattice().nelx);
      maxe = std::max(maxe, level(l).lattice().nelx);
    }
    mu_.alloc(maxe);
    mumax_.alloc(levels());
  }
  {
    const long long nd = 3 * level(levels() - 1).lattice().numnp;
    // A grid that cannot be halved further (an odd dimension) is also solved
    // exactly up to 1.5 dense_dofs: at 160x80x80 its 10x5x5 coarsest (1188
    // DOFs) smoothed took 283 PCG iterations and 119 ms per solve, solved
    // exactly 188 and 62 ms.  Beyond that the per-density factorization
    // outweighs the iterations saved (15x5x5, 1728 DOFs, at 60x20x20).
    const bool blocked = !halves(level(levels() - 1).problem().spec()) &&
                         (int)lv_.size() + 1 < std::max(1, o_.max_levels);
    dense_ = nd <= dense_max ||
             (blocked && nd <= std::min<long long>(MG_DENSE_MAX, dense_max + dense_max / 2));
    if (dense_) {
      ainv_.alloc(mg_dense_store_size(nd));
      anext_.alloc(mg_dense_store_size(nd));
      dpart_.alloc(mg_dense_part_size(level(levels() - 1).lattice()));
      HIPCHK(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&ev_side_, hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&ev_fact_, hipEventDisableTiming));
      for (int k = 0; k < 2; ++k) {
        info_[k].alloc(1);
        HIPCHK(hipHostMalloc((void **)&h_info_[k], sizeof(int), hipHostMallocDefault));
        *h_info_[k] = 0;
        if (rocblas_create_handle(&blas_[k]) != rocblas_status_success)
          throw Error(kInternal, "rocblas_create_handle failed");
        if (rocblas_set_stream(blas_[k], k == 0 ? f_.stream() : side_) != rocblas_status_success)
          throw Error(kInternal, "rocblas_set_stream failed");
        // library kernels that accumulate with atomics give run-to-run
        // differences; the solve must repeat bit for bit
        if (rocblas_set_atomics_mode(blas_[k], rocblas_atomics_not_allowed) !=
            rocblas_status_success)
          throw Error(kInternal, "rocblas_set_atomics_mode failed");
      }
    }
  }
  // W-cycle depth: automatic (w_levels < 0) by the fine mesh size, as measured
  // on the SIMP loops: none below 500,000 elements (the coarse levels' launches
  // cost more than the iterations saved), the first coarse level up to
  // 2,000,000, every coarse level above
  if (o_.w_levels >= 0) {
    wlev_ = o_.w_levels;
  } else {
    const long long ne = f_.lattice().nelx;
    wlev_ = ne >= 2000000 ? std::max(0, levels() - 2) : (ne >= 500000 ? 1 : 0);
  }
  z_.alloc(f_.n3_);
  y0_.alloc(f_.n3_);
  r0_.alloc(f_.n3_);
  if (o_.compensated) xlo_.alloc(f_.n3_);
  hipStream_t s = f_.stream();
  HIPCHK(hipMemsetAsync(z_.p, 0, z_.bytes(), s));
  HIPCHK(hipMemsetAsync(y0_.p, 0, y0_.bytes(), s));
  HIPCHK(hipMemsetAsync(r0_.p, 0, r0_.bytes(), s));
  for (int l = 1; l < levels(); ++l) {
    Context &c = level(l);
    for (DevBuf<double> *v : {&c.d_b, &c.d_x, &c.d_y, &c.d_r})
      HIPCHK(hipMemsetAsync(v->p, 0, v->bytes(), s));
  }
  f_.sync();
  zres_ = cycle(0, nullptr, false);
}

Multigrid::~Multigrid() {
  if (side_) (void)hipStreamSynchronize(side_);
  drop_graphs();
  for (int k = 0; k < 2; ++k) {
    if (h_info_[k]) (void)hipHostFree(h_info_[k]);
    if (blas_[k]) (void)rocblas_destroy_handle(blas_[k]);
  }
  if (ev_side_) (void)hipEventDestroy(ev_side_);
  if (ev_fact_) (void)hipEventDestroy(ev_fact_);
  if (side_) (void)hipStreamDestroy(side_);
  for (hipEvent_t e : ev_)
    if (e) (void)hipEventDestroy(e);
}

void Multigrid::drop_graphs() {
  for (int m = 0; m < 2; ++m)
    for (int p = 0; p < 2; ++p) {
      if (exec_[m][p]) (void)hipGraphExecDestroy(exec_[m][p]);
      if (graph_[m][p]) (void)hipGraphDestroy(graph_[m][p]);
      exec_[m][p] = nullptr;
      graph_[m][p] = nullptr;
    }
  mc_drop_graphs();
}

Context &Multigrid::level(int l) { return l == 0 ? f_ : *lv_[l - 1]; }

void Multigrid::refresh() {
  hipStream_t s = f_.stream();
  const GridSpec &g = f_.problem().spec();
  for (int l = 1; l < levels(); ++l) {
    Context &fc = level(l - 1), &cc = level(l);
    if (galerkin_) {
      // R A P element by element from the level above, then the damped
      // Jacobi factor theta / (mu_max diag)
      launch_gal_build(fc.lattice(), cc.lattice(), l == 1 ? fc.d_ke.p : kmat_[l - 1].p,
                       l == 1 ? fc.d_young.p : nullptr, fc.d_free, cc.d_free, kmat_[l], mu_, s);
      launch_gal_dinv(cc.lattice(), kmat_[l], cc.d_free, mu_, mumax_.p + l, o_.omega, cc.d_dinv, s);
      continue;
    }
    if (o_.coarse_avg == 2) {
      // the SIMP law of the mean density
      launch_mg_young(fc.lattice(), cc.lattice(), fc.d_dens, cc.d_dens, s, 0);
      launch_young(cc.d_dens, cc.d_young, cc.lattice().nelx, g.young1, g.young2, g.penalty, s);
    } else {
      launch_mg_young(fc.lattice(), cc.lattice(), fc.d_young, cc.d_young, s, o_.coarse_avg);
    }
    launch_diag_inv(cc.lattice(), cc.d_ke, cc.d_young, cc.d_free, cc.d_dinv, nullptr, s);
  }
  if (dense_) {
    if (!have_inverse_ || !overlap_) {
      factor_checked(ainv_);
      have_inverse_ = true;
      swap_pending_ = false;
    } else {
      // factored on the side stream while the first iterations run with the
      // previous density's inverse; swapped in at a fixed batch (before_batch)
      HIPCHK(hipEventRecord(ev_side_, s));
      HIPCHK(hipStreamWaitEvent(side_, ev_side_, 0));
      factor_into(anext_, 1, 0.0);
      HIPCHK(hipEventRecord(ev_fact_, side_));
      swap_pending_ = true;
    }
  }
}

// A = K of the coarsest grid (SPD on the free DOFs; constrained DOFs carry
// identity rows), plus shift * diag(A) on the free DOFs when shift > 0, then
// dst = A^-1 by Cholesky, on the fine context's stream (side = 0) or the side
// stream (side = 1), each with its own rocBLAS handle and status buffers.  The
// factorization's status goes to the pinned h_info_[side] (read by
// factor_checked and swap_inverse before the inverse is used).
void Multigrid::factor_into(double *dst, int side, double shift) {
  Context &cc = level(levels() - 1);
  hipStream_t s = side ? side_ : f_.stream();
  const int nd = (int)(3 * cc.lattice().numnp);
  launch_mg_dense_assemble(cc.lattice(), cc.d_ke, cc.d_young, cc.d_free, dst, s,
                           galerkin_ ? kmat_[levels() - 1].p : nullptr);
  if (shift > 0.0) launch_mg_dense_shift(nd, cc.lattice(), cc.d_free, shift, dst, s);
  auto chk = [](rocblas_status st, const char *what) {
    if (st != rocblas_status_success)
      throw Error(kInternal, std::string(what) + " failed: " + rocblas_status_to_string(st));
  };
  rocblas_handle hb = blas_[side];
  int *info = info_[side].p;
  chk(rocsolver_dpotrf(hb, rocblas_fill_lower, nd, dst, nd, info), "rocsolver_dpotrf");
  HIPCHK(hipMemcpyAsync(h_info_[side], info, sizeof(int), hipMemcpyDeviceToHost, s));
  chk(rocsolver_dpotri(hb, rocblas_fill_lower, nd, dst, nd, info), "rocsolver_dpotri");
  launch_mg_dense_pack(nd, dst, s);
  ++n_fact_;
}

// Synchronous factorization into dst on the fine context's stream, checked
// before the inverse is used.  A Cholesky breakdown (a non-positive pivot in
// floating point) is redone once with the diagonal of the free DOFs raised by
// a relative 1e-10 and counted in shifted().  The coarsest operator of a
// near-void field can be that ill-conditioned: its smallest eigenvalue can sit
// below the n eps ||A|| backward error of the factorization (see the CPU test
// of the 12x4x4 coarsest matrix).
void Multigrid::factor_checked(double *dst) {
  factor_into(dst, 0, 0.0);
  HIPCHK(hipStreamSynchronize(f_.stream()));
  if (*h_info_[0] == 0) return;
  ++n_shifted_;
  factor_into(dst, 0, 1e-10);
  HIPCHK(hipStreamSynchronize(f_.stream()));
  check_factor();
}

void Multigrid::swap_inverse(hipStream_t s) {
  if (!swap_pending_) return;
  swap_pending_ = false;
  // the side stream's status is read before its inverse replaces the current one
  HIPCHK(hipEventSynchronize(ev_fact_));
  if (*h_info_[1] != 0) {
    HIPCHK(hipStreamWaitEvent(f_.stream(), ev_fact_, 0));
    factor_checked(ainv_);
    if (s != f_.stream()) {
      HIPCHK(hipEventRecord(ev_side_, f_.stream()));
      HIPCHK(hipStreamWaitEvent(s, ev_side_, 0));
    }
    return;
  }
  HIPCHK(hipStreamWaitEvent(s, ev_fact_, 0));
  HIPCHK(hipMemcpyAsync(ainv_.p, anext_.p, ainv_.bytes(), hipMemcpyDeviceToDevice, s));
}

void Multigrid::before_batch(hipStream_t s) {
  if (batches_ == kSwapAfter && swap_pending_) {
    swap_inverse(s);
    // the Krylov directions so far used the previous preconditioner: restart
    // from z = M(r) with the new one (beta = 0)
    enqueue_restart();
  }
  // first phase over (the W-cycle, the FP32 fine level): from here on the
  // FP64 V-cycle, restarted from z = M_V(r).  The FP32 fine level ends at
  // kF32Switch x tol at the latest: its sweeps cannot resolve the last digits
  // of a field at its FP64 floor.  w_switch 0: the W-cycle runs to the end.
  const PcgScalars *h = f_.h_pcg.p;
  const double ws = f32cur_ ? std::max(o_.w_switch, kF32Switch) : o_.w_switch;
  if (wmode_ && !h->done && h->rr < ws * ws * tol_ * tol_ * h->bnorm2) {
    wmode_ = false;
    wcur_ = 0;
    f32cur_ = false;
    enqueue_restart();
  }
  ++batches_;
}

void Multigrid::finish() {
  swap_inverse(f_.stream());
  f_.sync();
  check_factor();
}

// V-cycle on level l; every launch on the fine context's stream.  Smoothing
// ping-pongs between the level's x and y buffers where the fused sweep runs
// (the z-marching path smooths x in place, with y holding K.x).
bool Multigrid::starts_swept(int l) const {
  const bool coarsest = l == levels() - 1;
  if (coarsest) return !dense_;
  return o_.nu_pre > 0;
}

void Multigrid::coarse_op(int l, int mode, const double *x, const double *b, double w,
                          double *out, hipStream_t s) {
  Context &c = level(l);
  if (galerkin_ && l > 0) {
    launch_gal_apply(mode == MG_JACOBI ? GAL_JACOBI : GAL_RESIDUAL, c.lattice(), kmat_[l], c.d_free,
                     x, b, c.d_dinv, out, s);
    return;
  }
  launch_mg_direct(mode, c.lattice(), c.kc_, c.d_young, c.d_free, x, b,
                   mode == MG_JACOBI ? c.d_dinv.p : nullptr, w, out, s);
}

bool Multigrid::fine_fp32_capable() const {
  return levels() > 1 && !direct_[0] && o_.nu_pre == 1 && o_.nu_post == 1;
}

void Multigrid::set_fine_fp32(bool on) {
  if (on == fp32_) return;
  if (on) {
    if (o_.nu_pre != 1 || o_.nu_post != 1)
      throw Error(kInvalidArgument, "the FP32 fine level needs nu_pre = nu_post = 1");
    if (!z32_.p) {
      z32_.alloc(f_.n3_);
      HIPCHK(hipMemsetAsync(z32_.p, 0, z32_.bytes(), f_.stream()));
      f_.sync();
    }
  }
  fp32_ = on;
  drop_graphs();
}

double *Multigrid::cycle(int l, const double *b, bool emit, bool x0_ready, ZDot *zd) {
  hipStream_t s = f_.stream();
  Context &c = level(l);
  const Lattice &L = c.lattice();
  const long long n = c.n3_;
  double *x = l == 0 ? z_.p : c.d_x.p;
  double *y = l == 0 ? y0_.p : c.d_y.p;
  double *r = l == 0 ? r0_.p : c.d_r.p;
  const double w = omega_[l];
  if (l == 0 && fine_fp32()) {
    // FP32 fine level: x0 = omega D^-1 b (written by the update when
    // x0_ready), the residual b - K x0 restricted with K x0 in FP32, the
    // coarse correction added in FP32, and the last sweep with its FP64
    // epilogue into the result buffer (with the z.r and z.q partials)
    if (zres_ && zres_ != y0_.p) throw Error(kInternal, "FP32 fine level: unexpected result buffer");
    Context &cc = level(1);
    if (emit) {
      if (!x0_ready) launch_mg_jacobi0_f32(n, w, c.d_dinv, b, z32_, s);
      // K x0 in FP64 storage (r0_): the restriction then reads the same
      // types as on the FP64 path
      launch_apply_k32(L, c.kc_, c.d_young32, c.d_free, z32_, r0_, s);
      const double *dinv_c = starts_swept(1) ? cc.d_dinv.p : nullptr;
      launch_mg_restrict(L, cc.lattice(), b, r0_, cc.d_free, cc.d_b, dinv_c, omega_[1], cc.d_x, s);
    }
    const double *xc = cycle(1, cc.d_b, emit);
    if (emit) {
      launch_mg_prolong_add_f32(L, cc.lattice(), xc, c.d_free, z32_.p, s);
      double *part = zd ? zd->part : c.d_partials.p;
      const double *q = zd ? zd->q : c.d_q.p;
      const unsigned cnt = launch_apply_k32_jacobi(L, c.kc_, c.d_young32, c.d_free, z32_, y0_, b,
                                                   c.d_dinv, w, q, part, s);
      if (zd) zd->count = cnt;
    }
    return y0_.p;
  }
  if (l == levels() - 1 && dense_) {
    if (emit) launch_mg_dense_solve(L, ainv_, b, x, dpart_, s);
    return x;
  }
  double *cur = x, *oth = y;
  auto sweep = [&]() {
    if (direct_[l]) {
      if (emit) coarse_op(l, MG_JACOBI, cur, b, w, oth, s);
      std::swap(cur, oth);
    } else if (c.kc_.kernel == 0) {
      // z-marching K.u with the sweep in its epilogue, out of place
      if (emit)
        launch_apply_k64_jacobi(L, c.kc_, c.d_young, c.d_free, cur, oth, b, c.d_dinv, w, nullptr,
                                nullptr, s);
      std::swap(cur, oth);
    } else if (emit) {
      launch_apply_k64(L, c.kc_, c.d_young, c.d_free, cur, oth, s);
      launch_mg_jacobi(n, w, c.d_dinv, b, oth, cur, s);
    }
  };
  // the first sweep from zero (levels >= 1: written by the restriction)
  if (l == 0 && emit && !x0_ready && starts_swept(0)) launch_mg_jacobi(n, w, c.d_dinv, b, nullptr, x, s);
  if (l == levels() - 1) {
    for (int i = 1; i < std::max(1, o_.nu_coarse); ++i) sweep();
    return cur;
  }
  const int pre = std::max(0, o_.nu_pre);
  for (int i = 1; i < pre; ++i) sweep();
  Context &cc = level(l + 1);
  if (emit) {
    // the coarse level's first sweep from zero comes with its right-hand side
    const double *dinv_c = starts_swept(l + 1) ? cc.d_dinv.p : nullptr;
    if (pre == 0) {
      // no pre-smoothing: x = 0, the residual is b
      launch_mg_restrict(L, cc.lattice(), b, nullptr, cc.d_free, cc.d_b, dinv_c, omega_[l + 1],
                         cc.d_x, s);
    } else if (direct_[l]) {
      coarse_op(l, MG_RESIDUAL, cur, b, 0.0, r, s);
      launch_mg_restrict(L, cc.lattice(), r, nullptr, cc.d_free, cc.d_b, dinv_c, omega_[l + 1],
                         cc.d_x, s);
    } else {
      // residual b - K x formed inside the restriction
      launch_apply_k64(L, c.kc_, c.d_young, c.d_free, cur, oth, s);
      launch_mg_restrict(L, cc.lattice(), b, oth, cc.d_free, cc.d_b, dinv_c, omega_[l + 1], cc.d_x, s);
    }
  }
  const double *xc = cycle(l + 1, cc.d_b, emit);
  // the second coarse visit's correction (W-cycle), added by the prolongation
  const double *xc2 = nullptr;
  if (l + 1 <= wcur_ && l + 1 < levels() - 1) {
    // W-cycle on this coarse level: a second coarse cycle on the remaining
    // coarse residual b_c - K_c xc; the prolongation takes xc (kept in d_u)
    // plus its correction
    const Lattice &Lc = cc.lattice();
    const long long nc = cc.n3_;
    double *x0 = starts_swept(l + 1) ? cc.d_x.p : nullptr;
    if (emit) {
      if (direct_[l + 1] && !galerkin_ && xc != cc.d_x.p) {
        // the copy, the residual and the first sweep from zero in one pass
        launch_mg_direct_wres(Lc, cc.kc_, cc.d_young, cc.d_free, xc, cc.d_b, cc.d_dinv,
                              omega_[l + 1], cc.d_u, cc.d_nv, x0, s);
      } else if (!direct_[l + 1] && cc.kc_.kernel == 0 && xc != cc.d_x.p) {
        // the same in the z-marching K.u's epilogue
        launch_apply_k64_wres(Lc, cc.kc_, cc.d_young, cc.d_free, xc, cc.d_b, cc.d_dinv,
                              omega_[l + 1], cc.d_u, cc.d_nv, x0, s);
      } else {
        launch_copy64(nc, xc, cc.d_u, s);
        if (direct_[l + 1]) {
          coarse_op(l + 1, MG_RESIDUAL, xc, cc.d_b, 0.0, cc.d_nv, s);
        } else {
          launch_apply_k64(Lc, cc.kc_, cc.d_young, cc.d_free, xc, cc.d_nv, s);
          launch_axpby(nc, 1.0, cc.d_b, -1.0, cc.d_nv, s);
        }
        // the first sweep from zero (the restriction writes it for the first cycle)
        if (x0) launch_mg_jacobi(nc, omega_[l + 1], cc.d_dinv, cc.d_nv, nullptr, x0, s);
      }
    }
    xc2 = cycle(l + 1, cc.d_nv, emit);
    xc = cc.d_u;
  }
  const int post = std::max(0, o_.nu_post);
  // a direct level's first post-sweep reads x + P xc straight from the
  // correction (same arithmetic as the two kernels, one pass fewer)
  const bool fused = fuse_ && direct_[l] && !(galerkin_ && l > 0) && pre > 0 && post > 0;
  int first = 0;
  if (fused) {
    if (emit)
      launch_mg_prolong_jacobi(L, cc.lattice(), c.kc_, c.d_young, c.d_free, xc, cur, b, c.d_dinv, w,
                               oth, s, xc2);
    std::swap(cur, oth);
    first = 1;
  } else if (emit) {
    // x = 0 without pre-smoothing: the correction is the iterate
    launch_mg_prolong_add(L, cc.lattice(), xc, c.d_free, cur, s, pre > 0, xc2);
  }
  for (int i = first; i < post; ++i) {
    if (i == post - 1 && l == 0 && !direct_[l]) {
      // The last fine sweep runs out of place (the V-cycle's result is always
      // the other buffer, whatever the kernel shape) and also yields the
      // partials of z.r and z.q.  With the z-marching K.u the sweep is fused
      // into the K.u's epilogue; else K.x goes through r.
      if (emit) {
        double *part = zd ? zd->part : c.d_partials.p;
        const double *q = zd ? zd->q : c.d_q.p;
        unsigned cnt;
        if (c.kc_.kernel == 0) {
          cnt = launch_apply_k64_jacobi(L, c.kc_, c.d_young, c.d_free, cur, oth, b, c.d_dinv, w, q,
                                        part, s);
        } else {
          launch_apply_k64(L, c.kc_, c.d_young, c.d_free, cur, r, s);
          cnt = launch_mg_jacobi_dot(n, w, c.d_dinv, b, r, cur, oth, q, part, s);
        }
        if (zd) zd->count = cnt;
      }
      std::swap(cur, oth);
    } else {
      sweep();
    }
  }
  return cur;
}

void Multigrid::enqueue_batch(int par) {
  Context &c = f_;
  hipStream_t s = c.stream();
  const Lattice &L = c.lattice();
  // the fine level's first sweep comes out of the update unless the fine level
  // does not start with one (solved exactly, or no pre-smoothing)
  const bool x0_fused = starts_swept(0);
  for (int j = 0; j < std::max(1, o_.batch_iters); ++j) {
    double *p_old = c.d_p[(j + par) & 1], *p_new = c.d_p[(j + par + 1) & 1];
    // p = z + beta p_old, q = K p, p.q partials
    const unsigned gs = launch_pcg_stencil64(L, c.kc_, c.d_young, c.d_free, zres_, nullptr, p_old,
                                             p_new, c.d_q, c.d_pcg, c.d_partials, c.d_counter, s,
                                             false);
    // alpha, x and r, r.r partials, z0 = omega D^-1 r
    const unsigned gu =
        fine_fp32()
            ? launch_fpcg_update_alpha(c.n3_, c.d_x, c.d_r, p_new, c.d_q, c.d_pcg, c.d_partials,
                                       gs, c.d_partials2, c.d_dinv, omega_[0], z32_.p, s, xlo_.p)
            : launch_fpcg_update_alpha(c.n3_, c.d_x, c.d_r, p_new, c.d_q, c.d_pcg, c.d_partials,
                                       gs, c.d_partials2, c.d_dinv, omega_[0],
                                       x0_fused ? z_.p : nullptr, s, xlo_.p);
    ZDot zd{c.d_q, c.d_partials, 0};
    cycle(0, c.d_r, true, x0_fused, &zd);
    if (!zd.count) zd.count = launch_zdot_partials(c.n3_, zres_, c.d_r, c.d_q, c.d_partials, s);
    // stop tests, then rho and beta
    launch_fpcg_check_beta(c.d_partials2, gu, c.d_partials, zd.count, c.d_pcg, s);
  }
}

void Multigrid::launch_batch(int par) {
  build_graph();
  before_batch(f_.stream());
  if (f_.use_graphs_)
    HIPCHK(hipGraphLaunch(exec_for(par), f_.stream()));
  else
    enqueue_batch(par & 1);
}

// Captures the batches of the FP64 V-cycle and, with a first phase (w_levels
// > 0 or the FP32 fine level), of that phase's cycle (the cycle shape is fixed
// in a graph; the solve picks the graph of its phase).
void Multigrid::build_graph() {
  if (!f_.use_graphs_ || exec_[0][0]) return;
  hipStream_t s = f_.stream();
  const int wsave = wcur_;
  const bool fsave = f32cur_;
  for (int m = 0; m < (two_phase() ? 2 : 1); ++m) {
    wcur_ = m ? wlev_ : 0;
    f32cur_ = m == 1 && fp32_;
    for (int par = 0; par < 2; ++par) {
      HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        enqueue_batch(par);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        wcur_ = wsave;
        f32cur_ = fsave;
        throw;
      }
      HIPCHK(hipStreamEndCapture(s, &graph_[m][par]));
      HIPCHK(hipGraphInstantiate(&exec_[m][par], graph_[m][par], nullptr, nullptr, 0));
    }
  }
  wcur_ = wsave;
  f32cur_ = fsave;
}

// one V-cycle (the learned preconditioner's base), whatever phase a solve is in
const double *Multigrid::vcycle(const double *r) {
  const int wsave = wcur_;
  wcur_ = 0;
  const double *z = cycle(0, r, true);
  wcur_ = wsave;
  return z;
}

void Multigrid::refresh_now() {
  hipStream_t s = f_.stream();
  if (refreshed_ != f_.density_version_) {
    swap_inverse(s);
    refresh();
    refreshed_ = f_.density_version_;
  }
  swap_inverse(s);  // a side-stream factorization goes in now
}

void Multigrid::enqueue_restart() {
  Context &c = f_;
  hipStream_t s = c.stream();
  launch_fpcg_restart_flag(c.d_pcg, s);
  cycle(0, c.d_r, true);
  launch_zdot_beta(c.n3_, zres_, c.d_r, c.d_q, c.d_pcg, c.d_partials, s);
}

void Multigrid::start(double tol, int maxit, hipStream_t s_ext, bool warm) {
  Context &c = f_;
  hipStream_t s = c.stream();
  for (hipEvent_t &e : ev_)
    if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if (s_ext != s) {
    HIPCHK(hipEventRecord(ev_[0], s_ext));
    HIPCHK(hipStreamWaitEvent(s, ev_[0], 0));
  }
  batches_ = 0;
  wmode_ = two_phase();
  wcur_ = wmode_ ? wlev_ : 0;
  f32cur_ = wmode_ && fp32_;
  tol_ = tol;
  best_ = INFINITY;
  since_best_ = 0;
  verifications_ = 0;
  if (refreshed_ != c.density_version_) {
    // a factorization still pending belongs to an older density
    swap_inverse(s);
    refresh();
    refreshed_ = c.density_version_;
  }
  PcgScalars *h = c.h_pcg.p;
  PcgScalars s0{};
  s0.tol = tol;
  s0.maxit = maxit;
  *h = s0;
  HIPCHK(hipMemcpyAsync(c.d_pcg.p, h, sizeof(PcgScalars), hipMemcpyHostToDevice, s));
  // x = 0, r = b, bnorm2 = b.b; with a guess, r = b - K x and the stop test on it
  if (warm) HIPCHK(hipMemcpyAsync(c.d_nv.p, c.d_x.p, c.d_x.bytes(), hipMemcpyDeviceToDevice, s));
  launch_pcg_init64(c.n3_, c.d_b, c.d_x, c.d_r, c.d_p[0], c.d_dinv, c.d_pcg, c.d_partials,
                    c.d_counter, s);
  if (xlo_.p) HIPCHK(hipMemsetAsync(xlo_.p, 0, xlo_.bytes(), s));
  if (warm) {
    HIPCHK(hipMemcpyAsync(c.d_x.p, c.d_nv.p, c.d_x.bytes(), hipMemcpyDeviceToDevice, s));
    launch_apply_k64(c.lattice(), c.kc_, c.d_young, c.d_free, c.d_x, c.d_y, s);
    launch_replace64(c.n3_, c.d_b, c.d_y, c.d_r, c.d_dinv, c.d_pcg, c.d_partials, s);
  }
  // z = M(r), rho = z.r, beta = 0 (a zero right-hand side, or a guess that
  // already passes, is done)
  enqueue_restart();
  HIPCHK(hipMemcpyAsync(h, c.d_pcg.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
  if (s_ext != s) {
    HIPCHK(hipEventRecord(ev_[1], s));
    HIPCHK(hipStreamWaitEvent(s_ext, ev_[1], 0));
  }
}

bool Multigrid::after_verify() {
  PcgScalars *h = f_.h_pcg.p;
  if (h->done) return true;  // the verified residual passes the recursive target
  hipStream_t s = f_.stream();
  const double bnorm = std::sqrt(h->bnorm2);
  const double tr = bnorm > 0.0 ? std::sqrt(h->rr) / bnorm : 0.0;
  if (tr < tol_) {
    // passes the solve's tolerance (the recursive target is tighter)
    h->done = 1;
    h->status = 0;
    return true;
  }
  ++verifications_;
  if (tr < best_) {
    best_ = tr;
    since_best_ = 0;
    if (!xbest_.p) xbest_.alloc(f_.n3_);
    HIPCHK(hipMemcpyAsync(xbest_.p, f_.d_x.p, f_.d_x.bytes(), hipMemcpyDeviceToDevice, s));
  } else {
    ++since_best_;
  }
  if (since_best_ < kStallVerifications && verifications_ <= kMaxVerifications) {
    // The next phase stops on a tighter recursive target t.  A phase that
    // stopped at the target t0 verified at tr: the recursive and the true
    // residual differ by about g = sqrt(tr^2 - t0^2) (rounding, in
    // quadrature), so the next phase aims at sqrt(t^2 + g^2) = 0.9 tol, at
    // least 0.3 tol.
    const double t0 = h->tol;
    const double g2 = std::max(0.0, tr * tr - t0 * t0);
    const double t = std::min(t0, std::sqrt(std::max(0.81 * tol_ * tol_ - g2, 0.09 * tol_ * tol_)));
    if (t < t0) {
      h->tol = t;
      HIPCHK(hipMemcpyAsync(f_.d_pcg.p, h, sizeof(PcgScalars), hipMemcpyHostToDevice, s));
    }
    return false;
  }
  // stalled at the floor: the best verified iterate, reported as not converged
  HIPCHK(hipMemcpyAsync(f_.d_x.p, xbest_.p, f_.d_x.bytes(), hipMemcpyDeviceToDevice, s));
  if (xlo_.p) HIPCHK(hipMemsetAsync(xlo_.p, 0, xlo_.bytes(), s));
  h->done = 1;
  h->status = 1;
  return true;
}

// A W-cycle phase whose preconditioned residual stopped being a descent
// direction (z.r <= 0: rounding in the repeated coarse corrections of a
// near-void field) goes on with the V-cycle from the current iterate
// (restart, beta = 0) instead of ending the solve.
bool Multigrid::recover_w(hipStream_t s_ext) {
  PcgScalars *h = f_.h_pcg.p;
  if (!(wmode_ && h->done && h->status == 4)) return false;
  Context &c = f_;
  hipStream_t s = c.stream();
  if (s_ext != s) {
    HIPCHK(hipEventRecord(ev_[0], s_ext));
    HIPCHK(hipStreamWaitEvent(s, ev_[0], 0));
  }
  wmode_ = false;
  wcur_ = 0;
  f32cur_ = false;
  ++w_fallbacks_;
  h->done = 0;
  h->status = 0;
  HIPCHK(hipMemcpyAsync(c.d_pcg.p, h, sizeof(PcgScalars), hipMemcpyHostToDevice, s));
  enqueue_restart();
  HIPCHK(hipMemcpyAsync(h, c.d_pcg.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
  if (s_ext != s) {
    HIPCHK(hipEventRecord(ev_[1], s));
    HIPCHK(hipStreamWaitEvent(s_ext, ev_[1], 0));
  }
  return true;
}

void Multigrid::check_factor() const {
  if (dense_ && *h_info_[0] != 0)
    throw Error(kNumerical, "coarsest multigrid matrix is not positive definite (Cholesky pivot " +
                                std::to_string(*h_info_[0]) + ")");
}

void Multigrid::verify(hipStream_t s_ext) {
  Context &c = f_;
  hipStream_t s = c.stream();
  if (s_ext != s) {
    HIPCHK(hipEventRecord(ev_[0], s_ext));
    HIPCHK(hipStreamWaitEvent(s, ev_[0], 0));
  }
  // the FP64 true residual replaces the recursive one; if it fails the test
  // the iteration restarts from z = M(r) with beta = 0
  launch_apply_k64(c.lattice(), c.kc_, c.d_young, c.d_free, c.d_x, c.d_y, s);
  launch_replace64(c.n3_, c.d_b, c.d_y, c.d_r, c.d_dinv, c.d_pcg, c.d_partials, s);
  enqueue_restart();
  HIPCHK(hipMemcpyAsync(c.h_pcg.p, c.d_pcg.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
  if (s_ext != s) {
    HIPCHK(hipEventRecord(ev_[1], s));
    HIPCHK(hipStreamWaitEvent(s_ext, ev_[1], 0));
  }
}

void Multigrid::solve(const icarat_solve_opts &o, icarat_solve_stats &st, bool warm) {
  if (o.precision != ICARAT_HIP_FP64 && o.precision != ICARAT_HIP_FP32)
    throw Error(kInvalidArgument,
                "the multigrid preconditioner runs an FP64 or an FP32 fine level (no BF16)");
  set_fine_fp32(o.precision == ICARAT_HIP_FP32);
  Context &c = f_;
  hipStream_t s = c.stream();
  start(o.tol, o.maxit, s, warm);
  build_graph();
  c.sync();
  PcgScalars *h = c.h_pcg.p;
  const double bnorm = std::sqrt(h->bnorm2);
  const int batch = std::max(1, o_.batch_iters);
  int repl = 0;
  for (;;) {
    if (!h->done) {
      const long long max_batches = (long long)(o.maxit - h->iter) / batch + 3;
      for (long long b = 0; b < max_batches && !h->done; ++b) {
        before_batch(s);
        if (c.use_graphs_)
          HIPCHK(hipGraphLaunch(exec_for(h->iter), s));
        else
          enqueue_batch(h->iter & 1);
        HIPCHK(hipMemcpyAsync(h, c.d_pcg.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
        c.sync();
        // a harvesting caller snapshots (x, r, V(r)) at batch boundaries
        if (c.batch_hook_ && !h->done) c.batch_hook_();
      }
      if (!h->done) throw Error(kInternal, "multigrid PCG did not reach its stop test");
    }
    if (h->status != 0) {
      if (recover_w(s)) {
        c.sync();
        continue;
      }
      break;
    }
    // verified stop
    verify(s);
    c.sync();
    ++repl;
    if (after_verify()) break;
  }
  finish();
  st.fallback = 0;
  st.refine_steps = repl;
  st.iters = h->iter;
  if (h->done && (h->status == 2 || h->status == 4))
    throw Error(kNumerical, "multigrid PCG breakdown (the V-cycle is not positive definite here)");
  st.rec_relres = bnorm > 0.0 ? std::sqrt(h->rr) / bnorm : 0.0;
  st.status = (h->done && h->status == 0) || bnorm == 0.0 ? ICARAT_HIP_OK : ICARAT_HIP_NOT_CONVERGED;
}

}  // namespace icarat_hip
