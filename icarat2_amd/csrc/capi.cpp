// capi.cpp -- the extern "C" boundary (include/icarat_hip.h): argument checks,
// exception-to-status translation and the per-context error text.
#include <cstring>
#include <new>
#include <algorithm>
#include <string>
#include <vector>

#include "../../include/icarat_hip.h"
#include "batch.hpp"
#include "context.hpp"
#include "gcmma.hpp"
#include "learned.hpp"
#include "mma.hpp"
#include "mmam.hpp"
#include "multigrid.hpp"
#include "unet.hpp"

struct icarat_hip_ctx {
  icarat_hip::Context *c = nullptr;
  std::string err;
};

struct icarat_hip_batch {
  icarat_hip::Batch *b = nullptr;
  std::string err;
};

struct icarat_hip_mmasolver {
  icarat_hip::DeviceMMAm *obj = nullptr;
  std::string err;
};

struct icarat_hip_gcmmasolver {
  icarat_hip::DeviceGCMMAm *obj = nullptr;
  std::string err;
};

struct icarat_hip_unet {
  hipStream_t s = nullptr;
  icarat_hip::UNet *net = nullptr;
  icarat_hip::NetGrid g{0, 0, 0};
  icarat_hip::DevBuf<__bf16> x;
  icarat_hip::DevBuf<float> xf, out, t;
  std::string err;
};

namespace icarat_hip {
icarat_hip_ctx *wrap_context(Context *c) {
  icarat_hip_ctx *h = new icarat_hip_ctx();
  h->c = c;
  return h;
}
Context *unwrap_context(icarat_hip_ctx *h) { return h ? h->c : nullptr; }
void set_context_error(icarat_hip_ctx *h, const std::string &m) {
  if (h) h->err = m;
}
}  // namespace icarat_hip

namespace {

template <class F>
int guard(icarat_hip_ctx *h, F &&f) {
  if (!h || !h->c) return ICARAT_HIP_INVALID_ARGUMENT;
  try {
    f(*h->c);
    h->err.clear();
    return ICARAT_HIP_OK;
  } catch (const icarat_hip::Error &e) {
    h->err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    h->err = "host allocation failed";
    return ICARAT_HIP_OUT_OF_MEMORY;
  } catch (const std::exception &e) {
    h->err = e.what();
    return ICARAT_HIP_INTERNAL;
  }
}

int fail(icarat_hip_ctx *h, int code, const char *msg) {
  if (h) h->err = msg;
  return code;
}

// a solver object's handle (obj, err): exception-to-status translation
template <class H, class F>
int obj_guard(H *h, F &&f) {
  if (!h || !h->obj) return ICARAT_HIP_INVALID_ARGUMENT;
  try {
    f(*h->obj);
    h->err.clear();
    return ICARAT_HIP_OK;
  } catch (const icarat_hip::Error &e) {
    h->err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    h->err = "host allocation failed";
    return ICARAT_HIP_OUT_OF_MEMORY;
  } catch (const std::exception &e) {
    h->err = e.what();
    return ICARAT_HIP_INTERNAL;
  }
}

}  // namespace

extern "C" {

int icarat_hip_device_available(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n > 0 ? 1 : 0;
}

int icarat_hip_create(const icarat_grid_desc *d, icarat_hip_ctx **out) {
  if (!d || !out) return ICARAT_HIP_INVALID_ARGUMENT;
  *out = nullptr;
  icarat_hip_ctx *h = new (std::nothrow) icarat_hip_ctx();
  if (!h) return ICARAT_HIP_OUT_OF_MEMORY;
  try {
    icarat_hip::GridSpec g{d->nx, d->ny, d->nz, d->lx, d->ly, d->lz,
                           d->young1, d->young2, d->poisson, d->penalty};
    h->c = new icarat_hip::Context(g);
  } catch (const icarat_hip::Error &e) {
    delete h;
    return e.code;
  } catch (...) {
    delete h;
    return ICARAT_HIP_INTERNAL;
  }
  *out = h;
  return ICARAT_HIP_OK;
}

int icarat_hip_destroy(icarat_hip_ctx *h) {
  if (!h) return ICARAT_HIP_INVALID_ARGUMENT;
  delete h->c;
  delete h;
  return ICARAT_HIP_OK;
}

const char *icarat_hip_last_error(const icarat_hip_ctx *h) { return h ? h->err.c_str() : "null context"; }

int icarat_hip_counts(const icarat_hip_ctx *h, int64_t *numnp, int64_t *nelx, int64_t *neq,
                      int32_t *numeq) {
  if (!h || !h->c) return ICARAT_HIP_INVALID_ARGUMENT;
  const auto &L = h->c->lattice();
  if (numnp) *numnp = L.numnp;
  if (nelx) *nelx = L.nelx;
  if (neq) *neq = 3 * L.numnp;
  if (numeq) *numeq = h->c->problem().dof_ready() ? h->c->problem().numeq() : -1;
  return ICARAT_HIP_OK;
}

int icarat_hip_add_dirichlet(icarat_hip_ctx *h, const int32_t flag[6], const double val[6],
                             const double range[6]) {
  if (!flag || !val || !range) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) {
    int f[6];
    for (int i = 0; i < 6; ++i) f[i] = flag[i];
    c.problem().add_dirichlet(f, val, range);
  });
}

int icarat_hip_make_dof(icarat_hip_ctx *h, int32_t *numeq) {
  return guard(h, [&](icarat_hip::Context &c) {
    const int n = c.make_dof();
    if (numeq) *numeq = n;
  });
}

int icarat_hip_add_neumann(icarat_hip_ctx *h, const double val[3], const double range[6]) {
  if (!val || !range) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) {
    c.problem().add_neumann(val, range);
    c.mark_fext_dirty();
  });
}

int icarat_hip_new_load_case(icarat_hip_ctx *h, int32_t *index) {
  return guard(h, [&](icarat_hip::Context &c) {
    const int i = c.new_load_case();
    if (index) *index = i;
  });
}

int icarat_hip_select_load_case(icarat_hip_ctx *h, int32_t index) {
  return guard(h, [&](icarat_hip::Context &c) { c.select_load_case(index); });
}

int icarat_hip_load_case_count(const icarat_hip_ctx *hc, int32_t *n) {
  icarat_hip_ctx *h = const_cast<icarat_hip_ctx *>(hc);
  if (!n) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) { *n = c.problem().num_load_cases(); });
}

int icarat_hip_get_dof(const icarat_hip_ctx *h, int32_t *dof) {
  if (!h || !h->c || !dof) return ICARAT_HIP_INVALID_ARGUMENT;
  if (!h->c->problem().dof_ready()) return ICARAT_HIP_NOT_READY;
  const auto &d = h->c->problem().dof();
  std::memcpy(dof, d.data(), d.size() * sizeof(int32_t));
  return ICARAT_HIP_OK;
}

int icarat_hip_get_fext(const icarat_hip_ctx *h, double *f) {
  if (!h || !h->c || !f) return ICARAT_HIP_INVALID_ARGUMENT;
  if (!h->c->problem().dof_ready()) return ICARAT_HIP_NOT_READY;
  const auto &v = h->c->problem().fext();
  std::memcpy(f, v.data(), v.size() * sizeof(double));
  return ICARAT_HIP_OK;
}

int icarat_hip_set_fext(icarat_hip_ctx *h, const double *f) {
  if (!f) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) {
    if (!c.problem().dof_ready()) throw icarat_hip::Error(ICARAT_HIP_NOT_READY, "make_dof has not been called");
    auto &v = c.problem().fext();
    std::memcpy(v.data(), f, v.size() * sizeof(double));
    c.mark_fext_dirty();
  });
}

int icarat_hip_get_ke0(const icarat_hip_ctx *h, double *ke) {
  if (!h || !h->c || !ke) return ICARAT_HIP_INVALID_ARGUMENT;
  std::memcpy(ke, h->c->problem().ke0(), 576 * sizeof(double));
  return ICARAT_HIP_OK;
}

int icarat_hip_set_density(icarat_hip_ctx *h, const double *x, int64_t n) {
  if (!x) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) {
    if (n != c.lattice().nelx) throw icarat_hip::Error(ICARAT_HIP_INVALID_ARGUMENT, "density length != nelx");
    c.set_density(x);
  });
}

int icarat_hip_apply_K(icarat_hip_ctx *h, const double *u, double *y, int32_t precision) {
  if (!u || !y) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) { c.apply_k(u, y, precision); });
}

int icarat_hip_get_diag(icarat_hip_ctx *h, double *d) {
  if (!d) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) { c.get_diag(d); });
}

int icarat_hip_solve_cases(icarat_hip_ctx *h, const icarat_solve_opts *o, double *u,
                           icarat_solve_stats *st) {
  if (!o) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null options");
  std::vector<icarat_solve_stats> s;
  int rc = guard(h, [&](icarat_hip::Context &c) {
    s.assign(c.problem().num_load_cases(), icarat_solve_stats{});
    c.solve_cases(*o, u, s.data());
  });
  if (st) std::copy(s.begin(), s.end(), st);
  if (rc == ICARAT_HIP_OK)
    for (const auto &x : s)
      if (x.status != ICARAT_HIP_OK) {
        h->err = "iteration cap reached before the tolerance";
        return x.status;
      }
  return rc;
}

int icarat_hip_solve(icarat_hip_ctx *h, const icarat_solve_opts *o, double *u,
                     icarat_solve_stats *st) {
  if (!o) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null options");
  icarat_solve_stats s{};
  int rc = guard(h, [&](icarat_hip::Context &c) { c.solve(*o, u, &s); });
  if (st) *st = s;
  if (rc == ICARAT_HIP_OK && s.status != ICARAT_HIP_OK) {
    h->err = "iteration cap reached before the tolerance";
    return s.status;
  }
  return rc;
}

int icarat_hip_true_residual(icarat_hip_ctx *h, const double *u, double *relres) {
  if (!u || !relres) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) { *relres = c.true_residual_host(u); });
}

int icarat_hip_compliance_sensitivity(icarat_hip_ctx *h, double vnorm, double *cc, double *dfds,
                                      double *dgds) {
  return guard(h, [&](icarat_hip::Context &c) { c.compliance_sensitivity(vnorm, cc, dfds, dgds); });
}

int icarat_hip_gcmma_set_params(icarat_hip_ctx *h, double move, double albefa, double asyminit,
                                double asymdec, double asyminc) {
  return guard(h, [&](icarat_hip::Context &c) {
    c.gcmma().set_params(move, albefa, asyminit, asymdec, asyminc);
  });
}

int icarat_hip_gcmma_reset(icarat_hip_ctx *h) {
  return guard(h, [&](icarat_hip::Context &c) { c.gcmma().reset(); });
}

int icarat_hip_gcmma_outer(icarat_hip_ctx *h, double *xmma, const double *xval, double f0x,
                           const double *df0dx, double fx, const double *dfdx, double xmin,
                           double xmax) {
  if (!xmma || !xval || !df0dx || !dfdx) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) {
    c.gcmma_host(0, xmma, 0.0, 0.0, xval, f0x, df0dx, fx, dfdx, xmin, xmax);
  });
}

int icarat_hip_gcmma_inner(icarat_hip_ctx *h, double *xmma, double f0xnew, double fxnew,
                           const double *xval, double f0x, const double *df0dx, double fx,
                           const double *dfdx, double xmin, double xmax) {
  if (!xmma || !xval || !df0dx || !dfdx) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) {
    c.gcmma_host(1, xmma, f0xnew, fxnew, xval, f0x, df0dx, fx, dfdx, xmin, xmax);
  });
}

int icarat_hip_gcmma_concheck(icarat_hip_ctx *h, double f0xnew, double fxnew,
                              int32_t *conservative) {
  if (!conservative) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) {
    *conservative = c.gcmma().concheck(f0xnew, fxnew) ? 1 : 0;
  });
}

int icarat_hip_mises(icarat_hip_ctx *h, double *mises, double *stress) {
  if (!mises) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) { c.mises_host(mises, stress); });
}

int icarat_hip_mises_noalloc(icarat_hip_ctx *h, double *mises, double *stress) {
  if (!mises) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) { c.mises_noalloc(mises, stress); });
}

int icarat_hip_filter_setup(icarat_hip_ctx *h, double radius) {
  return guard(h, [&](icarat_hip::Context &c) { c.filter_setup(radius); });
}

int icarat_hip_filter(icarat_hip_ctx *h, const double *in, double *out, int32_t is_sens) {
  if (!in || !out) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) { c.filter_host(in, out, is_sens ? 1 : 0); });
}

int icarat_hip_sensitivity_filter(icarat_hip_ctx *h, const double *x, const double *dfds,
                                  double *out) {
  if (!x || !dfds || !out) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) { c.sensitivity_filter_host(x, dfds, out); });
}

int icarat_hip_helm_setup(icarat_hip_ctx *h, double radius) {
  return guard(h, [&](icarat_hip::Context &c) { c.helm_setup(radius); });
}

int icarat_hip_helm_filter(icarat_hip_ctx *h, const double *in, double *out, int32_t is_sens) {
  if (!in || !out) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) { c.helm_host(in, out, is_sens ? 1 : 0); });
}

int icarat_hip_threshold(icarat_hip_ctx *h, const double *in, double beta, double T, double *out,
                         int32_t deriv) {
  if (!in || !out) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) { c.threshold_host(in, beta, T, out, deriv ? 1 : 0); });
}

int icarat_hip_oc_update(icarat_hip_ctx *h, double *x, const double *dfdx, double g,
                         const double *dgdx, double xmin, double xmax, int32_t use_projection,
                         double beta, int32_t *steps) {
  if (!x || !dfdx || !dgdx) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) {
    const int s = c.oc_host(x, dfdx, g, dgdx, xmin, xmax, use_projection, beta);
    if (steps) *steps = s;
  });
}

int icarat_hip_mma_set_params(icarat_hip_ctx *h, double move, double albefa, double asyminit,
                              double asymdec, double asyminc) {
  return guard(h, [&](icarat_hip::Context &c) {
    c.ensure_device();
    if (!c.mma_) c.mma_.reset(new icarat_hip::DeviceMMA(c));
    c.mma_->set_params(move, albefa, asyminit, asymdec, asyminc);
  });
}

int icarat_hip_mma_reset(icarat_hip_ctx *h) {
  return guard(h, [&](icarat_hip::Context &c) {
    c.ensure_device();
    if (!c.mma_) c.mma_.reset(new icarat_hip::DeviceMMA(c));
    c.mma_->reset();
    c.sync();
  });
}

int icarat_hip_mma_update(icarat_hip_ctx *h, double *x, const double *dfdx, double g,
                          const double *dgdx, double xmin, double xmax, int32_t *outer) {
  if (!x || !dfdx || !dgdx) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) {
    c.ensure_device();
    if (!c.mma_) c.mma_.reset(new icarat_hip::DeviceMMA(c));
    const long long n = c.lattice().nelx;
    hipStream_t s = c.stream();
    // element work vectors 5..7 (not used by the design loop's MMA path here)
    HIPCHK(hipMemcpyAsync(c.d_e[5].p, x, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c.d_e[6].p, dfdx, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c.d_e[4].p, dgdx, n * sizeof(double), hipMemcpyHostToDevice, s));
    const int o = c.mma_->update(c.d_e[5], c.d_e[6], g, c.d_e[4], xmin, xmax);
    HIPCHK(hipMemcpyAsync(x, c.d_e[5].p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    c.sync();
    if (outer) *outer = o;
  });
}

int icarat_hip_simp_run(icarat_hip_ctx *h, const icarat_simp_opts *o, double *x, double *hist,
                        icarat_simp_stats *st) {
  if (!o || !x) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) {
    icarat_simp_stats s{};
    c.simp_run(*o, x, hist, &s);
    if (st) *st = s;
  });
}

int icarat_hip_set_stencil(icarat_hip_ctx *h, int32_t kernel, int32_t sym_max, int32_t *sym_used) {
  if (kernel < 0 || kernel > 1 || sym_max < 0 || sym_max > 2)
    return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "kernel must be 0 or 1 and sym_max 0..2");
  return guard(h, [&](icarat_hip::Context &c) {
    c.set_stencil(kernel, sym_max);
    if (sym_used) *sym_used = c.kc_.sym;
  });
}

int icarat_hip_time_kernel(icarat_hip_ctx *h, int32_t kernel, int32_t reps, double *avg_ms,
                           double *flops, double *bytes) {
  return guard(h, [&](icarat_hip::Context &c) { c.time_kernel(kernel, reps, avg_ms, flops, bytes); });
}

// ---- learned preconditioner ------------------------------------------------
int icarat_hip_learned_default_opts(icarat_learned_opts *o) {
  if (!o) return ICARAT_HIP_INVALID_ARGUMENT;
  *o = icarat_learned_opts{};
  o->train_iters = 10;
  o->train_steps = 400;
  o->snapshot_every = 64;
  o->max_samples = 160;
  o->batch_iters = 8;
  o->stall_iters = 400;
  o->lr = 1e-3f;
  o->gate = 0.8f;
  o->seed = 1;
  o->base = 0;
  o->correction_scale = 1.f;
  return ICARAT_HIP_OK;
}

int icarat_hip_learned_setup(icarat_hip_ctx *h, const icarat_learned_opts *o) {
  icarat_learned_opts d;
  icarat_hip_learned_default_opts(&d);
  return guard(h, [&](icarat_hip::Context &c) { c.learned_setup(o ? *o : d); });
}

int icarat_hip_learned_harvest(icarat_hip_ctx *h, const icarat_solve_opts *o, double *u,
                               icarat_solve_stats *st) {
  if (!o) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null options");
  icarat_solve_stats s{};
  int rc = guard(h, [&](icarat_hip::Context &c) { c.learned_harvest(*o, u, &s); });
  if (st) *st = s;
  if (rc == ICARAT_HIP_OK && s.status != ICARAT_HIP_OK) {
    h->err = "iteration cap reached before the tolerance";
    return s.status;
  }
  return rc;
}

int icarat_hip_learned_train(icarat_hip_ctx *h, int32_t steps, double *loss) {
  if (steps < 0) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "steps must be >= 0");
  return guard(h, [&](icarat_hip::Context &c) {
    const double l = c.learned().train(steps);
    if (loss) *loss = l;
  });
}

int icarat_hip_learned_info(const icarat_hip_ctx *hc, icarat_learned_info *info) {
  icarat_hip_ctx *h = const_cast<icarat_hip_ctx *>(hc);
  if (!info) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) {
    icarat_hip::LearnedPrecond &lp = c.learned();
    icarat_learned_info d{};
    d.samples = lp.samples();
    d.active = lp.active() ? 1 : 0;
    d.steps = lp.total_steps();
    d.loss = lp.last_loss();
    d.val_loss = lp.val_loss();
    d.gain = lp.gain();
    d.n_params = icarat_hip::UNet::host_param_count();
    d.grid[0] = lp.grid().X;
    d.grid[1] = lp.grid().Y;
    d.grid[2] = lp.grid().Z;
    d.harvested = lp.harvested();
    d.fallback_iter = lp.fallback_iter();
    d.trial_iters = lp.trial_iters();
    d.trial_base_iters = lp.trial_base_iters();
    *info = d;
  });
}

int icarat_hip_learned_get_params(const icarat_hip_ctx *hc, float *params) {
  icarat_hip_ctx *h = const_cast<icarat_hip_ctx *>(hc);
  if (!params) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) { c.learned().net().get_params(params); });
}

int icarat_hip_learned_set_params(icarat_hip_ctx *h, const float *params) {
  if (!params) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) { c.learned().set_params(params); });
}

int icarat_hip_learned_apply(icarat_hip_ctx *h, const double *r, double *z) {
  if (!r || !z) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) { c.learned_apply(r, z); });
}

// ---- multigrid preconditioner -------------------------------------------------
int icarat_hip_mg_default_opts(icarat_mg_opts *o) {
  if (!o) return ICARAT_HIP_INVALID_ARGUMENT;
  *o = icarat_hip::mg_default_opts();
  return ICARAT_HIP_OK;
}

int icarat_hip_mg_setup(icarat_hip_ctx *h, const icarat_mg_opts *o) {
  icarat_mg_opts d;
  icarat_hip_mg_default_opts(&d);
  return guard(h, [&](icarat_hip::Context &c) { c.mg_setup(o ? *o : d); });
}

int icarat_hip_mg_info(const icarat_hip_ctx *hc, int32_t *levels, int32_t *dims) {
  icarat_hip_ctx *h = const_cast<icarat_hip_ctx *>(hc);
  return guard(h, [&](icarat_hip::Context &c) {
    if (!c.mg_) throw icarat_hip::Error(ICARAT_HIP_NOT_READY, "no multigrid hierarchy");
    const int n = c.mg_->levels();
    if (levels) *levels = n;
    if (dims)
      for (int l = 0; l < n; ++l) {
        const icarat_hip::Lattice &L = c.mg_->level(l).lattice();
        dims[3 * l] = L.nx;
        dims[3 * l + 1] = L.ny;
        dims[3 * l + 2] = L.nz;
      }
  });
}

int icarat_hip_mg_status(const icarat_hip_ctx *hc, icarat_mg_status *st) {
  icarat_hip_ctx *h = const_cast<icarat_hip_ctx *>(hc);
  if (!st) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) {
    if (!c.mg_) throw icarat_hip::Error(ICARAT_HIP_NOT_READY, "no multigrid hierarchy");
    icarat_mg_status o{};
    o.levels = c.mg_->levels();
    o.dense = c.mg_->dense() ? 1 : 0;
    o.direct_fine = c.mg_->direct(0) ? 1 : 0;
    o.fine_fp32 = c.mg_->fine_fp32_on() ? 1 : 0;
    o.factorizations = c.mg_->factorizations();
    o.shifted = c.mg_->shifted();
    o.w_levels = c.mg_->w_levels();
    o.w_fallbacks = c.mg_->w_fallbacks();
    *st = o;
  });
}

}  // extern "C"

// ---- the network alone ------------------------------------------------------
namespace {
template <class F>
int unet_guard(icarat_hip_unet *u, F &&f) {
  if (!u || !u->net) return ICARAT_HIP_INVALID_ARGUMENT;
  try {
    f(*u);
    u->err.clear();
    return ICARAT_HIP_OK;
  } catch (const icarat_hip::Error &e) {
    u->err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    u->err = "host allocation failed";
    return ICARAT_HIP_OUT_OF_MEMORY;
  } catch (const std::exception &e) {
    u->err = e.what();
    return ICARAT_HIP_INTERNAL;
  }
}

// host x[v*4 + c] -> device BF16 x[v*8 + c] (channels 4..7 zero)
void unet_upload(icarat_hip_unet &u, const float *x) {
  const long long nv = u.g.nvox();
  std::vector<float> x8(nv * 8, 0.f);
  for (long long v = 0; v < nv; ++v)
    for (int c = 0; c < 4; ++c) x8[v * 8 + c] = x[v * 4 + c];
  HIPCHK(hipMemcpyAsync(u.xf.p, x8.data(), nv * 8 * sizeof(float), hipMemcpyHostToDevice, u.s));
  icarat_hip::launch_f32_to_bf16(nv * 8, u.xf, u.x, u.s);
}

double unet_loss(icarat_hip_unet &u) {
  double l2[2];
  HIPCHK(hipMemcpyAsync(l2, u.net->loss2(), 2 * sizeof(double), hipMemcpyDeviceToHost, u.s));
  HIPCHK(hipStreamSynchronize(u.s));
  return l2[1] > 0.0 ? 0.5 * l2[0] / l2[1] : 0.0;
}
}  // namespace

extern "C" {

int icarat_hip_unet_create(int32_t X, int32_t Y, int32_t Z, uint64_t seed, icarat_hip_unet **out) {
  if (!out) return ICARAT_HIP_INVALID_ARGUMENT;
  *out = nullptr;
  if (X <= 0 || Y <= 0 || Z <= 0 || X % 4 || Y % 4 || Z % 4) return ICARAT_HIP_INVALID_ARGUMENT;
  icarat_hip_unet *u = new (std::nothrow) icarat_hip_unet();
  if (!u) return ICARAT_HIP_OUT_OF_MEMORY;
  try {
    HIPCHK(hipStreamCreateWithFlags(&u->s, hipStreamNonBlocking));
    u->g = icarat_hip::NetGrid{X, Y, Z};
    const long long nv = u->g.nvox();
    u->net = new icarat_hip::UNet(u->g, u->s, seed);
    u->x.alloc(nv * 8);
    u->xf.alloc(nv * 8);
    u->out.alloc(nv * 8);
    u->t.alloc(nv * 3);
  } catch (const icarat_hip::Error &e) {
    delete u->net;
    if (u->s) (void)hipStreamDestroy(u->s);
    delete u;
    return e.code;
  } catch (...) {
    delete u->net;
    if (u->s) (void)hipStreamDestroy(u->s);
    delete u;
    return ICARAT_HIP_INTERNAL;
  }
  *out = u;
  return ICARAT_HIP_OK;
}

int icarat_hip_unet_destroy(icarat_hip_unet *u) {
  if (!u) return ICARAT_HIP_INVALID_ARGUMENT;
  if (u->s) (void)hipStreamSynchronize(u->s);
  delete u->net;
  u->x.release();
  u->xf.release();
  u->out.release();
  u->t.release();
  if (u->s) (void)hipStreamDestroy(u->s);
  delete u;
  return ICARAT_HIP_OK;
}

const char *icarat_hip_unet_last_error(const icarat_hip_unet *u) { return u ? u->err.c_str() : "null network"; }

int icarat_hip_unet_param_count(int64_t *n) {
  if (!n) return ICARAT_HIP_INVALID_ARGUMENT;
  *n = icarat_hip::UNet::host_param_count();
  return ICARAT_HIP_OK;
}

int icarat_hip_unet_set_params(icarat_hip_unet *u, const float *p) {
  if (!p) return ICARAT_HIP_INVALID_ARGUMENT;
  return unet_guard(u, [&](icarat_hip_unet &n) { n.net->set_params(p); });
}

int icarat_hip_unet_get_params(const icarat_hip_unet *uc, float *p) {
  if (!p) return ICARAT_HIP_INVALID_ARGUMENT;
  icarat_hip_unet *u = const_cast<icarat_hip_unet *>(uc);
  return unet_guard(u, [&](icarat_hip_unet &n) { n.net->get_params(p); });
}

int icarat_hip_unet_forward(icarat_hip_unet *u, const float *x, float *out) {
  if (!x || !out) return ICARAT_HIP_INVALID_ARGUMENT;
  return unet_guard(u, [&](icarat_hip_unet &n) {
    const long long nv = n.g.nvox();
    unet_upload(n, x);
    n.net->forward(n.x, n.out);
    std::vector<float> o8(nv * 8);
    HIPCHK(hipMemcpyAsync(o8.data(), n.out.p, nv * 8 * sizeof(float), hipMemcpyDeviceToHost, n.s));
    HIPCHK(hipStreamSynchronize(n.s));
    for (long long v = 0; v < nv; ++v)
      for (int c = 0; c < 3; ++c) out[v * 3 + c] = o8[v * 8 + c];
  });
}

int icarat_hip_unet_grads(icarat_hip_unet *u, const float *x, const float *target, double *loss,
                          float *grads) {
  if (!x || !target) return ICARAT_HIP_INVALID_ARGUMENT;
  return unet_guard(u, [&](icarat_hip_unet &n) {
    unet_upload(n, x);
    HIPCHK(hipMemcpyAsync(n.t.p, target, n.g.nvox() * 3 * sizeof(float), hipMemcpyHostToDevice, n.s));
    n.net->backward_only(n.x, n.t);
    const double l = unet_loss(n);
    if (loss) *loss = l;
    if (grads) n.net->get_grads(grads);
  });
}

int icarat_hip_unet_train_step(icarat_hip_unet *u, const float *x, const float *target, float lr,
                               double *loss) {
  if (!x || !target || !(lr > 0.f)) return ICARAT_HIP_INVALID_ARGUMENT;
  return unet_guard(u, [&](icarat_hip_unet &n) {
    unet_upload(n, x);
    HIPCHK(hipMemcpyAsync(n.t.p, target, n.g.nvox() * 3 * sizeof(float), hipMemcpyHostToDevice, n.s));
    n.net->train_step(n.x, n.t, lr);
    const double l = unet_loss(n);
    if (loss) *loss = l;
  });
}

// ---- batched problems ---------------------------------------------------------
}  // extern "C"

namespace {
template <class F>
int batch_guard(icarat_hip_batch *b, F &&f) {
  if (!b || !b->b) return ICARAT_HIP_INVALID_ARGUMENT;
  try {
    f(*b->b);
    b->err.clear();
    return ICARAT_HIP_OK;
  } catch (const icarat_hip::Error &e) {
    b->err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    b->err = "host allocation failed";
    return ICARAT_HIP_OUT_OF_MEMORY;
  } catch (const std::exception &e) {
    b->err = e.what();
    return ICARAT_HIP_INTERNAL;
  }
}
}  // namespace

extern "C" {

int icarat_hip_batch_create(icarat_hip_batch **out) {
  if (!out) return ICARAT_HIP_INVALID_ARGUMENT;
  *out = nullptr;
  icarat_hip_batch *h = new (std::nothrow) icarat_hip_batch();
  if (!h) return ICARAT_HIP_OUT_OF_MEMORY;
  try {
    h->b = new icarat_hip::Batch();
  } catch (const icarat_hip::Error &e) {
    delete h;
    return e.code;
  } catch (...) {
    delete h;
    return ICARAT_HIP_INTERNAL;
  }
  *out = h;
  return ICARAT_HIP_OK;
}

int icarat_hip_batch_destroy(icarat_hip_batch *h) {
  if (!h) return ICARAT_HIP_INVALID_ARGUMENT;
  delete h->b;
  delete h;
  return ICARAT_HIP_OK;
}

const char *icarat_hip_batch_last_error(const icarat_hip_batch *h) {
  return h ? h->err.c_str() : "null batch";
}

int icarat_hip_batch_add(icarat_hip_batch *h, icarat_hip_ctx *ctx) {
  if (!ctx || !ctx->c) return ICARAT_HIP_INVALID_ARGUMENT;
  return batch_guard(h, [&](icarat_hip::Batch &b) { b.add(ctx->c); });
}

int icarat_hip_batch_set_budget(icarat_hip_batch *h, int64_t elements) {
  if (elements < 0) return ICARAT_HIP_INVALID_ARGUMENT;
  return batch_guard(h, [&](icarat_hip::Batch &b) { b.set_budget(elements); });
}

int icarat_hip_batch_info(const icarat_hip_batch *hc, int32_t *n, int64_t *launches) {
  icarat_hip_batch *h = const_cast<icarat_hip_batch *>(hc);
  return batch_guard(h, [&](icarat_hip::Batch &b) {
    if (n) *n = b.size();
    if (launches) *launches = b.graph_launches();
  });
}

int icarat_hip_batch_solve(icarat_hip_batch *h, const icarat_solve_opts *o,
                           icarat_solve_stats *stats) {
  if (!o) return ICARAT_HIP_INVALID_ARGUMENT;
  int cap = 0;
  int rc = batch_guard(h, [&](icarat_hip::Batch &b) {
    std::vector<icarat_solve_stats> s(b.size());
    b.solve(*o, s.data());
    for (int i = 0; i < b.size(); ++i) {
      if (s[i].status != ICARAT_HIP_OK) cap = 1;
      if (stats) stats[i] = s[i];
    }
  });
  if (rc == ICARAT_HIP_OK && cap) {
    h->err = "iteration cap reached before the tolerance in at least one problem";
    return ICARAT_HIP_NOT_CONVERGED;
  }
  return rc;
}

int icarat_hip_batch_simp_run(icarat_hip_batch *h, const icarat_simp_opts *o, double *const *x,
                              double *const *hist, icarat_simp_stats *stats) {
  if (!o || !x) return ICARAT_HIP_INVALID_ARGUMENT;
  return batch_guard(h, [&](icarat_hip::Batch &b) { b.simp_run(*o, x, hist, stats); });
}

int icarat_hip_batch_autotune(icarat_hip_batch *h, int32_t reps, int32_t *config, double *ms) {
  return batch_guard(h, [&](icarat_hip::Batch &b) {
    const int c = b.autotune(reps, ms);
    if (config) *config = c;
  });
}

int icarat_hip_mmasolver_create(int32_t n, int32_t m, double a, double c, double d,
                                icarat_hip_mmasolver **out) {
  if (!out) return ICARAT_HIP_INVALID_ARGUMENT;
  *out = nullptr;
  if (n < 1 || m < 1 || m > icarat_hip::kMmaMaxM) return ICARAT_HIP_INVALID_ARGUMENT;
  icarat_hip_mmasolver *h = new (std::nothrow) icarat_hip_mmasolver();
  if (!h) return ICARAT_HIP_OUT_OF_MEMORY;
  try {
    h->obj = new icarat_hip::DeviceMMAm(n, m, a, c, d);
  } catch (const icarat_hip::Error &e) {
    delete h;
    return e.code;
  } catch (...) {
    delete h;
    return ICARAT_HIP_INTERNAL;
  }
  *out = h;
  return ICARAT_HIP_OK;
}

int icarat_hip_mmasolver_destroy(icarat_hip_mmasolver *h) {
  if (!h) return ICARAT_HIP_INVALID_ARGUMENT;
  delete h->obj;
  delete h;
  return ICARAT_HIP_OK;
}

const char *icarat_hip_mmasolver_last_error(const icarat_hip_mmasolver *h) {
  return h ? h->err.c_str() : "null MMA solver";
}

int icarat_hip_mmasolver_set_params(icarat_hip_mmasolver *h, double move, double albefa,
                                    double asyminit, double asymdec, double asyminc) {
  return obj_guard(h, [&](icarat_hip::DeviceMMAm &m) {
    m.set_params(move, albefa, asyminit, asymdec, asyminc);
  });
}

int icarat_hip_mmasolver_reset(icarat_hip_mmasolver *h) {
  return obj_guard(h, [&](icarat_hip::DeviceMMAm &m) { m.reset(); });
}

int icarat_hip_mmasolver_update(icarat_hip_mmasolver *h, double *x, const double *dfdx,
                                const double *gx, const double *dgdx, const double *xmin,
                                const double *xmax, int32_t *outer) {
  if (!x || !dfdx || !gx || !dgdx || !xmin || !xmax) {
    if (h) h->err = "null argument";
    return ICARAT_HIP_INVALID_ARGUMENT;
  }
  return obj_guard(h, [&](icarat_hip::DeviceMMAm &m) {
    const int o = m.update(x, dfdx, gx, dgdx, xmin, xmax);
    if (outer) *outer = o;
  });
}

int icarat_hip_mmasolver_asymptotes(icarat_hip_mmasolver *h, double *low, double *upp) {
  return obj_guard(h, [&](icarat_hip::DeviceMMAm &m) { m.asymptotes(low, upp); });
}

int icarat_hip_gcmmasolver_create(int32_t n, int32_t m, double a, double c, double d,
                                  icarat_hip_gcmmasolver **out) {
  if (!out) return ICARAT_HIP_INVALID_ARGUMENT;
  *out = nullptr;
  if (n < 1 || m < 1 || m > icarat_hip::kMmaMaxM) return ICARAT_HIP_INVALID_ARGUMENT;
  icarat_hip_gcmmasolver *h = new (std::nothrow) icarat_hip_gcmmasolver();
  if (!h) return ICARAT_HIP_OUT_OF_MEMORY;
  try {
    h->obj = new icarat_hip::DeviceGCMMAm(n, m, a, c, d);
  } catch (const icarat_hip::Error &e) {
    delete h;
    return e.code;
  } catch (...) {
    delete h;
    return ICARAT_HIP_INTERNAL;
  }
  *out = h;
  return ICARAT_HIP_OK;
}

int icarat_hip_gcmmasolver_destroy(icarat_hip_gcmmasolver *h) {
  if (!h) return ICARAT_HIP_INVALID_ARGUMENT;
  delete h->obj;
  delete h;
  return ICARAT_HIP_OK;
}

const char *icarat_hip_gcmmasolver_last_error(const icarat_hip_gcmmasolver *h) {
  return h ? h->err.c_str() : "null GCMMA solver";
}

int icarat_hip_gcmmasolver_set_params(icarat_hip_gcmmasolver *h, double move, double albefa,
                                      double asyminit, double asymdec, double asyminc) {
  return obj_guard(h, [&](icarat_hip::DeviceGCMMAm &g) {
    g.set_params(move, albefa, asyminit, asymdec, asyminc);
  });
}

int icarat_hip_gcmmasolver_reset(icarat_hip_gcmmasolver *h) {
  return obj_guard(h, [&](icarat_hip::DeviceGCMMAm &g) { g.reset(); });
}

int icarat_hip_gcmmasolver_outer(icarat_hip_gcmmasolver *h, double *xmma, const double *xval,
                                 double f0x, const double *df0dx, const double *fx,
                                 const double *dfdx, const double *xmin, const double *xmax) {
  if (!xmma || !xval || !df0dx || !fx || !dfdx || !xmin || !xmax) {
    if (h) h->err = "null argument";
    return ICARAT_HIP_INVALID_ARGUMENT;
  }
  return obj_guard(h, [&](icarat_hip::DeviceGCMMAm &g) {
    g.outer(xmma, xval, f0x, df0dx, fx, dfdx, xmin, xmax);
  });
}

int icarat_hip_gcmmasolver_inner(icarat_hip_gcmmasolver *h, double *xmma, double f0xnew,
                                 const double *fxnew, const double *xval, double f0x,
                                 const double *df0dx, const double *fx, const double *dfdx,
                                 const double *xmin, const double *xmax) {
  if (!xmma || !fxnew || !xval || !df0dx || !fx || !dfdx || !xmin || !xmax) {
    if (h) h->err = "null argument";
    return ICARAT_HIP_INVALID_ARGUMENT;
  }
  return obj_guard(h, [&](icarat_hip::DeviceGCMMAm &g) {
    g.inner(xmma, f0xnew, fxnew, xval, f0x, df0dx, fx, dfdx, xmin, xmax);
  });
}

int icarat_hip_gcmmasolver_concheck(icarat_hip_gcmmasolver *h, double f0xnew, const double *fxnew,
                                    int32_t *ok) {
  if (!fxnew || !ok) {
    if (h) h->err = "null argument";
    return ICARAT_HIP_INVALID_ARGUMENT;
  }
  return obj_guard(h, [&](icarat_hip::DeviceGCMMAm &g) { *ok = g.concheck(f0xnew, fxnew) ? 1 : 0; });
}

int icarat_hip_gcmmasolver_state(icarat_hip_gcmmasolver *h, double *out) {
  if (!out) {
    if (h) h->err = "null argument";
    return ICARAT_HIP_INVALID_ARGUMENT;
  }
  return obj_guard(h, [&](icarat_hip::DeviceGCMMAm &g) { g.state(out); });
}

int icarat_hip_gaussian_field(int32_t nx, int32_t ny, int32_t nz, double h, uint64_t seed,
                              int32_t nblob, double smin, double smax, double mean, double lo,
                              double hi, double *out, double *mean_out) {
  if (!out || nx <= 0 || ny <= 0 || nz <= 0 || nblob <= 0) return ICARAT_HIP_INVALID_ARGUMENT;
  const double m = icarat_hip::gaussian_field(nx, ny, nz, h, seed, nblob, smin, smax, mean, lo, hi, out);
  if (mean_out) *mean_out = m;
  return ICARAT_HIP_OK;
}

}  // extern "C"
