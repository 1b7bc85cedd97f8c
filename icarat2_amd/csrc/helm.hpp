// helm.hpp -- Helmholtz PDE filter on the voxel lattice (reference
// optimization/filter_helm.hpp), solved exactly by fast diagonalisation.
//
// On a box of axis-aligned voxels the assembled filter matrix
// sum_e (N N^T + c B^T B) jac (c = (r / 2 sqrt 3)^2, no boundary condition)
// is separable:
//   A = Mx (x) My (x) Mz + c (Kx (x) My (x) Mz + Mx (x) Ky (x) Mz + Mx (x) My (x) Kz)
// with the 1-D consistent mass M and stiffness K of linear elements along each
// axis (2-point Gauss integrates both exactly).  The generalised eigenvectors
// K V = M V L, V^T M V = I (one small problem per axis, rocSOLVER dsygvd)
// diagonalise it: A^-1 = (Vx (x) Vy (x) Vz) (I + c (Lx + Ly + Lz))^-1 (...)^T.
// One filter pass is then three FP64 GEMMs along x, y and z, a pointwise
// scaling, and the three transposed GEMMs: a direct solve, where the
// reference runs Jacobi-PCG to a relative residual of 1e-10.
#pragma once
#include "hip_util.hpp"
#include "helm_kernels.hpp"

struct _rocblas_handle;  // rocBLAS handle type (rocblas/rocblas.h)

namespace icarat_hip {

class Context;

class HelmFilter {
 public:
  HelmFilter(Context &c, double radius);
  ~HelmFilter();
  HelmFilter(const HelmFilter &) = delete;
  HelmFilter &operator=(const HelmFilter &) = delete;

  double radius() const { return r_; }
  // FilterHelmholtz::densityFilter(in, solver, log, is_sens) on device element
  // vectors, enqueued on the context's stream (in and out may alias)
  void apply(const double *in, double *out, int is_sens);
  // largest |V^T M V - I| over the three axes (checked at set-up)
  double orthogonality_error() const { return orth_err_; }

 private:
  void solve_inplace();  // a_ <- A^-1 a_ (b_ is scratch)

  Context &c_;
  double r_, coef_, vol_e_;
  int n_[3];                    // nodes per axis
  DevBuf<double> V_[3], lam_[3];
  DevBuf<double> a_, b_;        // node-lattice work vectors (numnp)
  ::_rocblas_handle *blas_ = nullptr;
  double orth_err_ = 0.0;
};

}  // namespace icarat_hip
