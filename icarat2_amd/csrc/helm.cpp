// helm.cpp -- see helm.hpp.
#include "helm.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "context.hpp"

namespace icarat_hip {

namespace {

void chk(rocblas_status st, const char *what) {
  if (st != rocblas_status_success)
    throw Error(kInternal, std::string(what) + " failed: " + rocblas_status_to_string(st));
}

// 1-D linear elements of length h on n nodes, natural boundary conditions:
// consistent mass M = h/6 [.. 1 4 1 ..] (h/3 at the ends) and stiffness
// K = 1/h [.. -1 2 -1 ..] (1/h at the ends), dense column-major
void axis_matrices(int n, double h, std::vector<double> &K, std::vector<double> &M) {
  K.assign((size_t)n * n, 0.0);
  M.assign((size_t)n * n, 0.0);
  for (int e = 0; e + 1 < n; ++e) {
    const int a = e, b = e + 1;
    K[a + (size_t)n * a] += 1.0 / h;
    K[b + (size_t)n * b] += 1.0 / h;
    K[a + (size_t)n * b] -= 1.0 / h;
    K[b + (size_t)n * a] -= 1.0 / h;
    M[a + (size_t)n * a] += h / 3.0;
    M[b + (size_t)n * b] += h / 3.0;
    M[a + (size_t)n * b] += h / 6.0;
    M[b + (size_t)n * a] += h / 6.0;
  }
}

}  // namespace

HelmFilter::HelmFilter(Context &c, double radius) : c_(c), r_(radius) {
  if (!(radius > 0.0)) throw Error(kInvalidArgument, "Helmholtz filter radius must be > 0");
  c_.ensure_device();
  const GridSpec &g = c_.problem().spec();
  const Lattice &L = c_.lattice();
  // reference filter_helm.hpp:201: IKe *= (r / (2 sqrt 3))^2
  coef_ = std::pow(radius / (2 * std::sqrt(3.0)), 2);
  const double h[3] = {g.lx / g.nx, g.ly / g.ny, g.lz / g.nz};
  vol_e_ = h[0] * h[1] * h[2];
  n_[0] = L.NX;
  n_[1] = L.NY;
  n_[2] = L.NZ;
  if ((long long)L.NX * L.NY > INT32_MAX || (long long)L.NY * L.NZ > INT32_MAX)
    throw Error(kInvalidArgument, "grid too large for the Helmholtz transforms");
  a_.alloc(L.numnp);
  b_.alloc(L.numnp);
  if (rocblas_create_handle(&blas_) != rocblas_status_success)
    throw Error(kInternal, "rocblas_create_handle failed");
  hipStream_t s = c_.stream();
  chk(rocblas_set_stream(blas_, s), "rocblas_set_stream");
  // no atomics in the library kernels: results repeat bit for bit
  chk(rocblas_set_atomics_mode(blas_, rocblas_atomics_not_allowed), "rocblas_set_atomics_mode");
  DevBuf<double> dM, dE;
  DevBuf<int> info(1);
  for (int d = 0; d < 3; ++d) {
    const int n = n_[d];
    std::vector<double> K, M;
    axis_matrices(n, h[d], K, M);
    V_[d].alloc((size_t)n * n);
    lam_[d].alloc(n);
    dM.alloc((size_t)n * n);
    dE.alloc(n);
    HIPCHK(hipMemcpyAsync(V_[d].p, K.data(), K.size() * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dM.p, M.data(), M.size() * sizeof(double), hipMemcpyHostToDevice, s));
    // K v = lambda M v; V overwrites K, normalised V^T M V = I
    chk(rocsolver_dsygvd(blas_, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper, n,
                         V_[d].p, n, dM.p, n, lam_[d].p, dE.p, info.p),
        "rocsolver_dsygvd");
    int h_info = 0;
    HIPCHK(hipMemcpyAsync(&h_info, info.p, sizeof(int), hipMemcpyDeviceToHost, s));
    std::vector<double> V((size_t)n * n), lam(n);
    HIPCHK(hipMemcpyAsync(V.data(), V_[d].p, V.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(lam.data(), lam_[d].p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (h_info != 0)
      throw Error(kNumerical, "Helmholtz filter: generalised eigenproblem failed (info " +
                                  std::to_string(h_info) + ")");
    // K >= 0 (one zero eigenvalue: the constants), M > 0
    for (int i = 0; i < n; ++i)
      if (!(lam[i] > -1e-10 * std::fabs(lam[n - 1])))
        throw Error(kNumerical, "Helmholtz filter: negative eigenvalue of the 1-D stiffness");
    // probe V^T M V = I with two fixed vectors: |V^T M V w - w| / |w|
    for (int t = 0; t < 2; ++t) {
      std::vector<double> w(n), vw(n, 0.0), mvw(n, 0.0), out(n, 0.0);
      for (int i = 0; i < n; ++i) w[i] = t == 0 ? 1.0 : std::cos(0.7 * i + 0.3);
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) vw[i] += V[i + (size_t)n * j] * w[j];
      for (int i = 0; i < n; ++i) {
        double s2 = M[i + (size_t)n * i] * vw[i];
        if (i > 0) s2 += M[i + (size_t)n * (i - 1)] * vw[i - 1];
        if (i + 1 < n) s2 += M[i + (size_t)n * (i + 1)] * vw[i + 1];
        mvw[i] = s2;
      }
      double num = 0.0, den = 0.0;
      for (int j = 0; j < n; ++j) {
        double s2 = 0.0;
        for (int i = 0; i < n; ++i) s2 += V[i + (size_t)n * j] * mvw[i];
        num += (s2 - w[j]) * (s2 - w[j]);
        den += w[j] * w[j];
      }
      orth_err_ = std::max(orth_err_, std::sqrt(num / den));
    }
  }
  if (!(orth_err_ < 1e-10))
    throw Error(kNumerical, "Helmholtz filter: eigenvectors not M-orthonormal");
}

HelmFilter::~HelmFilter() {
  if (blas_) (void)rocblas_destroy_handle(blas_);
}

// a_ <- (V (I + c L)^-1 V^T) a_ with V = Vx (x) Vy (x) Vz on the x-fastest lattice
void HelmFilter::solve_inplace() {
  const int NX = n_[0], NY = n_[1], NZ = n_[2];
  const double one = 1.0, zero = 0.0;
  const rocblas_operation N = rocblas_operation_none, T = rocblas_operation_transpose;
  chk(rocblas_set_stream(blas_, c_.stream()), "rocblas_set_stream");
  // x: (NX x NY*NZ) column-major, b = Vx^T a
  chk(rocblas_dgemm(blas_, T, N, NX, NY * NZ, NX, &one, V_[0].p, NX, a_.p, NX, &zero, b_.p, NX),
      "rocblas_dgemm (x)");
  // y: for every z plane (NX x NY), a_k = b_k Vy
  chk(rocblas_dgemm_strided_batched(blas_, N, N, NX, NY, NY, &one, b_.p, NX, (rocblas_stride)NX * NY,
                                    V_[1].p, NY, 0, &zero, a_.p, NX, (rocblas_stride)NX * NY, NZ),
      "rocblas_dgemm_strided_batched (y)");
  // z: (NX*NY x NZ), b = a Vz
  chk(rocblas_dgemm(blas_, N, N, NX * NY, NZ, NZ, &one, a_.p, NX * NY, V_[2].p, NZ, &zero, b_.p,
                    NX * NY),
      "rocblas_dgemm (z)");
  launch_helm_scale(c_.lattice(), lam_[0], lam_[1], lam_[2], coef_, b_, c_.stream());
  chk(rocblas_dgemm(blas_, N, T, NX * NY, NZ, NZ, &one, b_.p, NX * NY, V_[2].p, NZ, &zero, a_.p,
                    NX * NY),
      "rocblas_dgemm (z^T)");
  chk(rocblas_dgemm_strided_batched(blas_, N, T, NX, NY, NY, &one, a_.p, NX, (rocblas_stride)NX * NY,
                                    V_[1].p, NY, 0, &zero, b_.p, NX, (rocblas_stride)NX * NY, NZ),
      "rocblas_dgemm_strided_batched (y^T)");
  chk(rocblas_dgemm(blas_, N, N, NX, NY * NZ, NX, &one, V_[0].p, NX, b_.p, NX, &zero, a_.p, NX),
      "rocblas_dgemm (x^T)");
}

void HelmFilter::apply(const double *in, double *out, int is_sens) {
  const Lattice &L = c_.lattice();
  hipStream_t s = c_.stream();
  // element loads x_e V_e / 8 per node (/ 8 again for sensitivities,
  // filter_helm.hpp:218-221); read back sum_a val_a (/ 8 for densities, :252-253)
  launch_helm_rhs(L, in, is_sens ? vol_e_ / 64.0 : vol_e_ / 8.0, a_, s);
  solve_inplace();
  launch_helm_extract(L, a_, is_sens ? 1.0 : 0.125, out, s);
}

}  // namespace icarat_hip
