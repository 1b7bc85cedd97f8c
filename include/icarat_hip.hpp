// icarat_hip.hpp -- header-only C++ facade over the C ABI (icarat_hip.h) with
// the class names, argument meaning and error behaviour of ICARAT's operator
// API for the SIMP compliance path, so that a driver can switch by changing
// types:
//   LinearAnalysis::solve(solvName, log)     analysis/linear.hpp:55-104
//   Filter::densityFilter / threshold / thresholdDerivative
//                                            optimization/filter.hpp:21-45,106-155,216-248
//   FilterHelmholtz::densityFilter / threshold / thresholdDerivative
//                                            optimization/filter_helm.hpp:21-62,76-181
//   OC::UpdateVariables                      optimization/opt-solver/OCSolver.hpp:16-79
//   MMASolver (ctor, Set*, Update, Reset)    optimization/opt-solver/MMASolver.hpp:62-116
//   GCMMASolver (ctor, Set*, OuterUpdate, InnerUpdate, ConCheck, Reset)
//                                            optimization/opt-solver/GCMMASolver.hpp:62-85
// Like the reference, a fatal condition prints to std::cerr and calls
// exit(1); no exception crosses this header.  Mesh, material, boundary
// conditions and loads are held by Device (MeshIcarat + Elastic::setParameter +
// BCIcarat on a box, io/mesh_icarat.hpp:72-139, problem/elasticity.cpp:19-33,
// io/boundary_condition_icarat.hpp:72-410).
#ifndef ICARAT_HIP_HPP
#define ICARAT_HIP_HPP

#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "icarat_hip.h"

namespace icarat_hip {

inline void check(int rc, const icarat_hip_ctx *ctx, const char *what) {
  if (rc == ICARAT_HIP_OK) return;
  std::cerr << "Error : " << what << ": " << (ctx ? icarat_hip_last_error(ctx) : "")
            << " (status " << rc << ")" << std::endl;
  std::exit(1);
}

// Voxel box problem on the GPU: mesh, material, supports and loads.
class Device {
 public:
  Device(int nx, int ny, int nz, double lx, double ly, double lz, double young1, double young2,
         double poisson, double penalty) {
    icarat_grid_desc d{nx, ny, nz, lx, ly, lz, young1, young2, poisson, penalty};
    check(icarat_hip_create(&d, &ctx_), nullptr, "icarat_hip_create");
    int64_t np = 0, ne = 0, neq = 0;
    int32_t numeq = 0;
    icarat_hip_counts(ctx_, &np, &ne, &neq, &numeq);
    numnp_ = np;
    nelx_ = ne;
  }
  ~Device() { icarat_hip_destroy(ctx_); }
  Device(const Device &) = delete;
  Device &operator=(const Device &) = delete;

  // [[flag x6], [val x6], [[x0,x1],[y0,y1],[z0,z1]]] of a Dirichlet entry
  void addDirichlet(const int32_t flag[6], const double val[6], const double range[6]) {
    check(icarat_hip_add_dirichlet(ctx_, flag, val, range), ctx_, "add_dirichlet");
  }
  int makeDOF() {
    int32_t n = 0;
    check(icarat_hip_make_dof(ctx_, &n), ctx_, "makeDOF");
    numeq_ = n;
    return n;
  }
  // [[val x3], [[x0,x1],[y0,y1],[z0,z1]]] of a Neumann entry
  void addNeumann(const double val[3], const double range[6]) {
    check(icarat_hip_add_neumann(ctx_, val, range), ctx_, "add_neumann");
  }
  std::vector<int32_t> dof() const {
    std::vector<int32_t> d(3 * numnp_);
    check(icarat_hip_get_dof(ctx_, d.data()), ctx_, "dof");
    return d;
  }
  std::vector<double> fext() const {
    std::vector<double> f(numeq_);
    check(icarat_hip_get_fext(ctx_, f.data()), ctx_, "fext");
    return f;
  }
  // Elastic::setDesignS for every element: physical densities, nelx entries
  void setDesign(const std::vector<double> &xphys) {
    check(icarat_hip_set_density(ctx_, xphys.data(), (int64_t)xphys.size()), ctx_, "setDesign");
  }
  long long numnp() const { return numnp_; }
  long long nelx() const { return nelx_; }
  int numeq() const { return numeq_; }
  icarat_hip_ctx *handle() const { return ctx_; }

 private:
  icarat_hip_ctx *ctx_ = nullptr;
  long long numnp_ = 0, nelx_ = 0;
  int numeq_ = -1;
};

// LinearAnalysis with matrix-free solver names next to the reference's:
//   "HIP_MF_PCG"          FP64 Jacobi-PCG (Scaling_CG semantics, verified FP64 stop)
//   "HIP_MF_PCG_FP32"     FP32 MFMA contraction inside FP64 reliable updates
//   "HIP_MF_PCG_BF16"     BF16 MFMA contraction inside FP64 reliable updates
//   "HIP_MF_FPCG_LEARNED" learned preconditioner (icarat_hip_learned_setup first)
//   "HIP_MF_PCG_MG"       FP64 PCG preconditioned by a geometric multigrid V-cycle
//                         (hierarchy built on first use, icarat_hip_mg_setup to tune)
class LinearAnalysis {
 public:
  explicit LinearAnalysis(Device &dev, double tol = 1e-8, int maxIter = 10000)
      : dev_(dev), tol_(tol), maxit_(maxIter) {}
  void solve(const std::string &solvName, bool log) {
    icarat_solve_opts o{tol_, maxit_, ICARAT_HIP_FP64, ICARAT_HIP_PRECOND_JACOBI, 0.1, 200, 0};
    if (solvName == "HIP_MF_PCG_FP32") o.precision = ICARAT_HIP_FP32;
    else if (solvName == "HIP_MF_PCG_BF16") o.precision = ICARAT_HIP_BF16;
    else if (solvName == "HIP_MF_FPCG_LEARNED") o.precond = ICARAT_HIP_PRECOND_LEARNED;
    else if (solvName == "HIP_MF_PCG_MG") o.precond = ICARAT_HIP_PRECOND_MG;
    else if (solvName != "HIP_MF_PCG") {
      std::cerr << "Error : unknown solver name " << solvName << std::endl;
      std::exit(1);
    }
    val_.assign(dev_.numeq() > 0 ? dev_.numeq() : 0, 0.0);
    int rc = icarat_hip_solve(dev_.handle(), &o, val_.data(), &stats_);
    // like Scaling_CG, reaching the iteration cap is reported, not fatal
    if (rc != ICARAT_HIP_NOT_CONVERGED) check(rc, dev_.handle(), "solve");
    if (log)
      std::cout << "iteration: " << stats_.iters << "  residual norm (FP64 true): "
                << stats_.true_relres << std::endl;
  }
  const std::vector<double> &val() const { return val_; }        // Value::val (numeq)
  const icarat_solve_stats &stats() const { return stats_; }
  // compliance f.u and dfds / dgds of the driver's sensitivity()
  double compliance() {
    double c = 0.0;
    check(icarat_hip_compliance_sensitivity(dev_.handle(), 1.0, &c, nullptr, nullptr),
          dev_.handle(), "compliance");
    return c;
  }
  void sensitivity(double vnorm, std::vector<double> &dfds, std::vector<double> &dgds) {
    dfds.resize(dev_.nelx());
    dgds.resize(dev_.nelx());
    check(icarat_hip_compliance_sensitivity(dev_.handle(), vnorm, nullptr, dfds.data(),
                                            dgds.data()),
          dev_.handle(), "sensitivity");
  }

 private:
  Device &dev_;
  double tol_;
  int maxit_;
  std::vector<double> val_;
  icarat_solve_stats stats_{};
};

class Filter {
 public:
  Filter(Device &dev, double radius) : dev_(dev) {
    check(icarat_hip_filter_setup(dev_.handle(), radius), dev_.handle(), "Filter");
  }
  std::vector<double> densityFilter(const std::vector<double> &design_s, bool isSens) {
    std::vector<double> out(design_s.size());
    check(icarat_hip_filter(dev_.handle(), design_s.data(), out.data(), isSens ? 1 : 0),
          dev_.handle(), "densityFilter");
    return out;
  }
  // Filter::sensitivityFilter (optimization/filter.hpp:76-104)
  std::vector<double> sensitivityFilter(const std::vector<double> &design_s,
                                        const std::vector<double> &dfds) {
    std::vector<double> out(design_s.size());
    check(icarat_hip_sensitivity_filter(dev_.handle(), design_s.data(), dfds.data(), out.data()),
          dev_.handle(), "sensitivityFilter");
    return out;
  }
  std::vector<double> threshold(const std::vector<double> &design_s, double beta, double T) {
    std::vector<double> out(design_s.size());
    check(icarat_hip_threshold(dev_.handle(), design_s.data(), beta, T, out.data(), 0),
          dev_.handle(), "threshold");
    return out;
  }
  std::vector<double> thresholdDerivative(const std::vector<double> &design_s, double beta,
                                          double T) {
    std::vector<double> out(design_s.size());
    check(icarat_hip_threshold(dev_.handle(), design_s.data(), beta, T, out.data(), 1),
          dev_.handle(), "thresholdDerivative");
    return out;
  }

 private:
  Device &dev_;
};

// FilterHelmholtz (optimization/filter_helm.hpp:21-62): the solver name and log
// flag of densityFilter are accepted for signature compatibility; the nodal
// system is solved directly (fast diagonalisation) whatever the name.
class FilterHelmholtz {
 public:
  FilterHelmholtz(Device &dev, double radius) : dev_(dev) {
    check(icarat_hip_helm_setup(dev_.handle(), radius), dev_.handle(), "FilterHelmholtz");
  }
  std::vector<double> densityFilter(const std::vector<double> &design_s,
                                    const std::string & /*solvName*/, int /*log*/,
                                    bool sensFlag) {
    std::vector<double> out(design_s.size());
    check(icarat_hip_helm_filter(dev_.handle(), design_s.data(), out.data(), sensFlag ? 1 : 0),
          dev_.handle(), "densityFilter");
    return out;
  }
  std::vector<double> threshold(const std::vector<double> &design_s, double beta, double T) {
    std::vector<double> out(design_s.size());
    check(icarat_hip_threshold(dev_.handle(), design_s.data(), beta, T, out.data(), 0),
          dev_.handle(), "threshold");
    return out;
  }
  std::vector<double> thresholdDerivative(const std::vector<double> &design_s, double beta,
                                          double T) {
    std::vector<double> out(design_s.size());
    check(icarat_hip_threshold(dev_.handle(), design_s.data(), beta, T, out.data(), 1),
          dev_.handle(), "thresholdDerivative");
    return out;
  }

 private:
  Device &dev_;
};

// Bounds must be uniform (every reference driver passes Optimize(nelx, 1, xmin, xmax)).
inline bool uniform(const double *v, long long n) {
  for (long long i = 1; i < n; ++i)
    if (v[i] != v[0]) return false;
  return true;
}

// OC::UpdateVariables with the drivers' volume callback computed on the device:
// x' -> filter -> (projection) -> sum * element volume (opt_macro_oc.cpp:153-164).
class OC {
 public:
  OC(Device &dev, int n) : dev_(dev), n_(n) {}
  void UpdateVariables(std::vector<double> &xk, std::vector<double> &dfdx, double g,
                       std::vector<double> &dgdx, std::vector<double> &xmin,
                       std::vector<double> &xmax, bool useProjection = false,
                       double beta = 1.0) {
    if ((long long)xk.size() != n_ || !uniform(xmin.data(), n_) || !uniform(xmax.data(), n_)) {
      std::cerr << "Error : OC::UpdateVariables needs n entries and uniform bounds" << std::endl;
      std::exit(1);
    }
    int32_t steps = 0;
    check(icarat_hip_oc_update(dev_.handle(), xk.data(), dfdx.data(), g, dgdx.data(), xmin[0],
                               xmax[0], useProjection ? 1 : 0, beta, &steps),
          dev_.handle(), "OC::UpdateVariables");
  }

 private:
  Device &dev_;
  long long n_;
};

// MMASolver(n, m, a, c, d) on the device: any n, 1 <= m <= 8 constraints
// (icarat_hip_mmasolver_*; not tied to a Device)
class MMASolver {
 public:
  MMASolver(int n, int m, double a = 0.0, double c = 100000.0, double d = 0.0) {
    if (icarat_hip_mmasolver_create(n, m, a, c, d, &h_) != ICARAT_HIP_OK) {
      std::cerr << "Error : MMASolver on the device supports n >= 1 and 1 <= m <= 8" << std::endl;
      std::exit(1);
    }
  }
  MMASolver(Device &, int n, int m, double a = 0.0, double c = 100000.0, double d = 0.0)
      : MMASolver(n, m, a, c, d) {}
  MMASolver(const MMASolver &) = delete;
  MMASolver &operator=(const MMASolver &) = delete;
  ~MMASolver() {
    if (h_) icarat_hip_mmasolver_destroy(h_);
  }
  void SetAsymptotes(double init, double decrease, double increase) {
    asyminit_ = init;
    asymdec_ = decrease;
    asyminc_ = increase;
    push();
  }
  void SetMoveLimit(double m) {
    move_ = m;
    push();
  }
  void SetOscillation(double a) {
    albefa_ = a;
    push();
  }
  void ConstraintModification(bool) {}
  void Update(double *xval, const double *dfdx, const double *gx, const double *dgdx,
              const double *xmin, const double *xmax) {
    done(icarat_hip_mmasolver_update(h_, xval, dfdx, gx, dgdx, xmin, xmax, nullptr),
         "MMASolver::Update");
  }
  void Reset() { done(icarat_hip_mmasolver_reset(h_), "MMASolver::Reset"); }

 private:
  void push() {
    done(icarat_hip_mmasolver_set_params(h_, move_, albefa_, asyminit_, asymdec_, asyminc_),
         "MMASolver parameters");
  }
  void done(int rc, const char *what) {
    if (rc == ICARAT_HIP_OK) return;
    std::cerr << "Error : " << what << ": " << icarat_hip_mmasolver_last_error(h_) << " (status "
              << rc << ")" << std::endl;
    std::exit(1);
  }
  icarat_hip_mmasolver *h_ = nullptr;
  double move_ = 0.15, albefa_ = 0.2, asyminit_ = 0.2, asymdec_ = 0.7, asyminc_ = 1.2;
};

// GCMMASolver(n, m, a, c, d) on the device: any n, 1 <= m <= 8 constraints
// (optimization/opt-solver/GCMMASolver.hpp:62-85; icarat_hip_gcmmasolver_*, not
// tied to a Device)
class GCMMASolver {
 public:
  GCMMASolver(int n, int m, double a = 0.0, double c = 1000.0, double d = 0.0) {
    if (icarat_hip_gcmmasolver_create(n, m, a, c, d, &h_) != ICARAT_HIP_OK) {
      std::cerr << "Error : GCMMASolver on the device supports n >= 1 and 1 <= m <= 8"
                << std::endl;
      std::exit(1);
    }
  }
  GCMMASolver(Device &, int n, int m, double a = 0.0, double c = 1000.0, double d = 0.0)
      : GCMMASolver(n, m, a, c, d) {}
  GCMMASolver(const GCMMASolver &) = delete;
  GCMMASolver &operator=(const GCMMASolver &) = delete;
  ~GCMMASolver() {
    if (h_) icarat_hip_gcmmasolver_destroy(h_);
  }
  void SetAsymptotes(double init, double decrease, double increase) {
    asyminit_ = init;
    asymdec_ = decrease;
    asyminc_ = increase;
    push();
  }
  void SetMoveLimit(double m) {
    move_ = m;
    push();
  }
  void SetOscillation(double a) {
    albefa_ = a;
    push();
  }
  void OuterUpdate(double *xmma, const double *xval, double f0x, const double *df0dx,
                   const double *fx, const double *dfdx, const double *xmin,
                   const double *xmax) {
    done(icarat_hip_gcmmasolver_outer(h_, xmma, xval, f0x, df0dx, fx, dfdx, xmin, xmax),
         "GCMMASolver::OuterUpdate");
  }
  void InnerUpdate(double *xmma, double f0xnew, const double *fxnew, const double *xval,
                   double f0x, const double *df0dx, const double *fx, const double *dfdx,
                   const double *xmin, const double *xmax) {
    done(icarat_hip_gcmmasolver_inner(h_, xmma, f0xnew, fxnew, xval, f0x, df0dx, fx, dfdx, xmin,
                                      xmax),
         "GCMMASolver::InnerUpdate");
  }
  bool ConCheck(double f0xnew, const double *fxnew) {
    int32_t ok = 0;
    done(icarat_hip_gcmmasolver_concheck(h_, f0xnew, fxnew, &ok), "GCMMASolver::ConCheck");
    return ok != 0;
  }
  void Reset() { done(icarat_hip_gcmmasolver_reset(h_), "GCMMASolver::Reset"); }

 private:
  void push() {
    done(icarat_hip_gcmmasolver_set_params(h_, move_, albefa_, asyminit_, asymdec_, asyminc_),
         "GCMMASolver parameters");
  }
  void done(int rc, const char *what) {
    if (rc == ICARAT_HIP_OK) return;
    std::cerr << "Error : " << what << ": " << icarat_hip_gcmmasolver_last_error(h_)
              << " (status " << rc << ")" << std::endl;
    std::exit(1);
  }
  icarat_hip_gcmmasolver *h_ = nullptr;
  double move_ = 0.15, albefa_ = 0.2, asyminit_ = 0.2, asymdec_ = 0.7, asyminc_ = 1.2;
};

// addPlot(xx, yy, filename, isFirst) and exportGraph(yy, filename) /
// exportGraph(xx, yy, filename) of io/export_graph.hpp:15-83.
inline void addPlot(double xx, double yy, const std::string &filename, int isFirst) {
  if (icarat_hip_add_plot(filename.c_str(), xx, yy, isFirst) != ICARAT_HIP_OK) {
    std::cerr << "Error : cannot write " << filename << std::endl;
    std::exit(1);
  }
}
inline void exportGraph(const std::vector<double> &yy, const std::string &filename) {
  if (icarat_hip_export_graph(filename.c_str(), nullptr, yy.data(), (int64_t)yy.size()) !=
      ICARAT_HIP_OK) {
    std::cerr << "Error : cannot write " << filename << std::endl;
    std::exit(1);
  }
}
inline void exportGraph(const std::vector<double> &xx, const std::vector<double> &yy,
                        const std::string &filename) {
  if (xx.size() != yy.size() ||
      icarat_hip_export_graph(filename.c_str(), xx.data(), yy.data(), (int64_t)yy.size()) !=
          ICARAT_HIP_OK) {
    std::cerr << "Error : cannot write " << filename << std::endl;
    std::exit(1);
  }
}

}  // namespace icarat_hip

#endif  // ICARAT_HIP_HPP
