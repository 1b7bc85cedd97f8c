// ref_optim_wrap.cpp -- C entry points around the reference's own std-only
// optimizers, compiled from the reference sources where they lie
// (optimization/opt-solver/MMASolver.cpp, GCMMASolver.cpp, OCSolver.hpp,
// and the std-only headers optimization/opt_base.hpp, io/export_graph.hpp)
// into oracle/_ref/libref_optim.so by oracle/Makefile.
//
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle's OC and MMA
// and GCMMA restatements against the reference implementation on seeded inputs.
// Nothing in the product links or loads it.
#include <algorithm>
#include <cmath>
#include <iostream>
#include <vector>

#include "io/export_graph.hpp"
#include "optimization/opt-solver/GCMMASolver.hpp"
#include "optimization/opt-solver/MMASolver.hpp"
#include "optimization/opt-solver/OCSolver.hpp"
#include "optimization/opt_base.hpp"

extern "C" {

void *ref_mma_create(int n, int m, double a, double c, double d) {
  return new icarat::MMASolver(n, m, a, c, d);
}
void ref_mma_destroy(void *h) { delete (icarat::MMASolver *)h; }
void ref_mma_set(void *h, double move, double albefa, double asyminit, double asymdec,
                 double asyminc) {
  auto *M = (icarat::MMASolver *)h;
  M->SetMoveLimit(move);
  M->SetOscillation(albefa);
  M->SetAsymptotes(asyminit, asymdec, asyminc);
}
void ref_mma_update(void *h, double *x, const double *dfdx, const double *gx,
                    const double *dgdx, const double *xmin, const double *xmax) {
  ((icarat::MMASolver *)h)->Update(x, dfdx, gx, dgdx, xmin, xmax);
}

void *ref_gcmma_create(int n, int m, double a, double c, double d) {
  return new icarat::GCMMASolver(n, m, a, c, d);
}
void ref_gcmma_destroy(void *h) { delete (icarat::GCMMASolver *)h; }
void ref_gcmma_set(void *h, double move, double albefa, double asyminit, double asymdec,
                   double asyminc) {
  auto *G = (icarat::GCMMASolver *)h;
  G->SetMoveLimit(move);
  G->SetOscillation(albefa);
  G->SetAsymptotes(asyminit, asymdec, asyminc);
}
void ref_gcmma_outer(void *h, double *xmma, const double *xval, double f0x, const double *df0dx,
                     const double *fx, const double *dfdx, const double *xmin,
                     const double *xmax) {
  ((icarat::GCMMASolver *)h)->OuterUpdate(xmma, xval, f0x, df0dx, fx, dfdx, xmin, xmax);
}
void ref_gcmma_inner(void *h, double *xmma, double f0xnew, const double *fxnew,
                     const double *xval, double f0x, const double *df0dx, const double *fx,
                     const double *dfdx, const double *xmin, const double *xmax) {
  ((icarat::GCMMASolver *)h)->InnerUpdate(xmma, f0xnew, fxnew, xval, f0x, df0dx, fx, dfdx, xmin,
                                          xmax);
}
int ref_gcmma_concheck(void *h, double f0xnew, const double *fxnew) {
  return ((icarat::GCMMASolver *)h)->ConCheck(f0xnew, fxnew) ? 1 : 0;
}

typedef double (*ref_volume_fn)(const double *x, int n, void *user);

// OC::UpdateVariables with a caller-supplied volume function gkp1(x').
void ref_oc_update(int n, double *x, double *dfdx, double g, double *dgdx,
                   double *xmin, double *xmax, ref_volume_fn gk, void *user) {
  icarat::OC oc(n);
  std::vector<double> xv(x, x + n), df(dfdx, dfdx + n), dg(dgdx, dgdx + n);
  std::vector<double> lo(xmin, xmin + n), hi(xmax, xmax + n);
  oc.UpdateVariables(xv, df, g, dg, lo, hi,
                     [&](std::vector<double> &xk) { return gk(xk.data(), n, user); });
  std::copy(xv.begin(), xv.end(), x);
}

// addPlot(step, value, filename, step) as the drivers call it (an int step)
void ref_add_plot(const char *filename, int step, double value, int is_first) {
  icarat::addPlot(step, value, std::string(filename), is_first);
}
// exportGraph(yy, filename) and exportGraph(xx, yy, filename)
void ref_export_graph(const char *filename, const double *xx, const double *yy, int n) {
  std::vector<double> y(yy, yy + n);
  if (xx) {
    std::vector<double> x(xx, xx + n);
    icarat::exportGraph(x, y, std::string(filename));
  } else {
    icarat::exportGraph(y, std::string(filename));
  }
}
// Optimize::isConvergence(eps, minstep) over a compliance history f[0..n),
// optstep = k + 1 at entry k; out[k] = its return value (0: converged)
void ref_is_convergence(const double *f, int n, double eps, double minstep, int *out) {
  icarat::Optimize o(1, 1, 0.0, 1.0);
  for (int k = 0; k < n; ++k) {
    o.optstep = k + 1;
    o.object_f = f[k];
    std::streambuf *old = std::cout.rdbuf(nullptr);  // the banner it prints
    out[k] = o.isConvergence(eps, minstep);
    std::cout.rdbuf(old);
  }
}
}
