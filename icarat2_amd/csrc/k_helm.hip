// k_helm.hip -- the element/node transfers and the spectral scaling of the
// Helmholtz PDE filter on the voxel lattice (the dense 1-D transforms between
// them are FP64 GEMMs, see helm.cpp).
//
// With trilinear voxels the Gauss sums of the reference (optimization/
// filter_helm.hpp:191-257) reduce to closed forms: sum_ip N_a(ip) = 1 at
// every node, so int_e N_a = V_e / 8, the element load is x_e V_e / 8 per
// node (/ 8 again in the sensitivity mode) and the element value read back
// is sum_a val_a (/ 8 in the density mode).
#include <hip/hip_runtime.h>

#include "helm_kernels.hpp"

namespace icarat_hip {

namespace {

constexpr int NT = 256;

inline unsigned blocks_for(long long n, int nt) { return (unsigned)((n + nt - 1) / nt); }

// rhs[a] = s * sum of in[e] over the (up to 8) elements that share node a
__global__ __launch_bounds__(NT) void k_helm_rhs(Lattice L, const double *__restrict__ in,
                                                 double s, double *__restrict__ rhs) {
  const long long n = blockIdx.x * (long long)NT + threadIdx.x;
  if (n >= L.numnp) return;
  const int i = (int)(n % L.NX);
  const int j = (int)((n / L.NX) % L.NY);
  const int k = (int)(n / ((long long)L.NX * L.NY));
  double acc = 0.0;
  for (int ez = k - 1; ez <= k; ++ez) {
    if (ez < 0 || ez >= L.nz) continue;
    for (int ey = j - 1; ey <= j; ++ey) {
      if (ey < 0 || ey >= L.ny) continue;
      const long long row = (long long)L.nx * (ey + (long long)L.ny * ez);
      if (i - 1 >= 0) acc += in[row + i - 1];
      if (i < L.nx) acc += in[row + i];
    }
  }
  rhs[n] = s * acc;
}

// out[e] = s * sum of val over the 8 nodes of element e
__global__ __launch_bounds__(NT) void k_helm_extract(Lattice L, const double *__restrict__ val,
                                                     double s, double *__restrict__ out) {
  const long long e = blockIdx.x * (long long)NT + threadIdx.x;
  if (e >= L.nelx) return;
  const int i = (int)(e % L.nx);
  const int j = (int)((e / L.nx) % L.ny);
  const int k = (int)(e / ((long long)L.nx * L.ny));
  const long long sy = L.NX, sz = (long long)L.NX * L.NY;
  const long long b = i + sy * j + sz * k;
  const double bottom = (val[b] + val[b + 1]) + (val[b + sy] + val[b + sy + 1]);
  const double top = (val[b + sz] + val[b + sz + 1]) + (val[b + sz + sy] + val[b + sz + sy + 1]);
  out[e] = s * (bottom + top);
}

// v[i,j,k] /= 1 + c (lx[i] + ly[j] + lz[k])
__global__ __launch_bounds__(NT) void k_helm_scale(Lattice L, const double *__restrict__ lx,
                                                   const double *__restrict__ ly,
                                                   const double *__restrict__ lz, double c,
                                                   double *__restrict__ v) {
  const long long n = blockIdx.x * (long long)NT + threadIdx.x;
  if (n >= L.numnp) return;
  const int i = (int)(n % L.NX);
  const int j = (int)((n / L.NX) % L.NY);
  const int k = (int)(n / ((long long)L.NX * L.NY));
  v[n] = v[n] / (1.0 + c * (lx[i] + ly[j] + lz[k]));
}

}  // namespace

void launch_helm_rhs(const Lattice &L, const double *in, double s, double *rhs, hipStream_t st) {
  hipLaunchKernelGGL(k_helm_rhs, dim3(blocks_for(L.numnp, NT)), dim3(NT), 0, st, L, in, s, rhs);
  HIPCHK(hipGetLastError());
}

void launch_helm_extract(const Lattice &L, const double *val, double s, double *out,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_helm_extract, dim3(blocks_for(L.nelx, NT)), dim3(NT), 0, st, L, val, s,
                     out);
  HIPCHK(hipGetLastError());
}

void launch_helm_scale(const Lattice &L, const double *lx, const double *ly, const double *lz,
                       double c, double *v, hipStream_t st) {
  hipLaunchKernelGGL(k_helm_scale, dim3(blocks_for(L.numnp, NT)), dim3(NT), 0, st, L, lx, ly, lz,
                     c, v);
  HIPCHK(hipGetLastError());
}

}  // namespace icarat_hip
