// ref_optim_wrap.cpp -- C entry points around the reference's own std-only
// optimizers, compiled from the reference sources where they lie
// (optimization/opt-solver/MMASolver.cpp, GCMMASolver.cpp, OCSolver.hpp)
// into oracle/_ref/libref_optim.so by oracle/Makefile.
//
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle's OC and MMA
// and GCMMA restatements against the reference implementation on seeded inputs.
// Nothing in the product links or loads it.
#include <algorithm>
#include <cmath>
#include <iostream>
#include <vector>

#include "optimization/opt-solver/GCMMASolver.hpp"
#include "optimization/opt-solver/MMASolver.hpp"
#include "optimization/opt-solver/OCSolver.hpp"

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
}
