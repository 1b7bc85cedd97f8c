// capi.cpp -- the extern "C" boundary (include/icarat_hip.h): argument checks,
// exception-to-status translation and the per-context error text.
#include <cstring>
#include <new>
#include <string>

#include "../../include/icarat_hip.h"
#include "context.hpp"
#include "mma.hpp"

struct icarat_hip_ctx {
  icarat_hip::Context *c = nullptr;
  std::string err;
};

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

int icarat_hip_filter_setup(icarat_hip_ctx *h, double radius) {
  return guard(h, [&](icarat_hip::Context &c) { c.filter_setup(radius); });
}

int icarat_hip_filter(icarat_hip_ctx *h, const double *in, double *out, int32_t is_sens) {
  if (!in || !out) return fail(h, ICARAT_HIP_INVALID_ARGUMENT, "null argument");
  return guard(h, [&](icarat_hip::Context &c) { c.filter_host(in, out, is_sens ? 1 : 0); });
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

int icarat_hip_time_kernel(icarat_hip_ctx *h, int32_t kernel, int32_t reps, double *avg_ms,
                           double *flops, double *bytes) {
  return guard(h, [&](icarat_hip::Context &c) { c.time_kernel(kernel, reps, avg_ms, flops, bytes); });
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
