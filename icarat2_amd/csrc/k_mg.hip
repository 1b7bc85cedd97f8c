// k_mg.hip -- kernels of the geometric multigrid preconditioner: transfers
// between a node lattice and the lattice of the 2x coarser voxel grid, the
// coarse Young's moduli and the damped Jacobi sweep.
//
// Coarse node (I, J, K) sits on fine node (2I, 2J, 2K).  Prolongation is
// trilinear interpolation (1-D weights 1 at the coincident node, 1/2 at the
// two neighbours); restriction is its transpose.  Both are masked by the free
// DOFs of their output lattice.
#include <hip/hip_runtime.h>

#include "mg_kernels.hpp"

namespace icarat_hip {

namespace {

inline unsigned blocks_for(long long n, int nt) { return (unsigned)((n + nt - 1) / nt); }

__device__ __forceinline__ double w1(int d) { return d == 0 ? 1.0 : 0.5; }

// bc[coarse] = sum over the 27 fine nodes around 2I of w * rf[fine], free DOFs only
__global__ void k_mg_restrict(Lattice Lf, Lattice Lc, const double *__restrict__ rf,
                              const uint8_t *__restrict__ fmask_c, double *__restrict__ bc) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= Lc.numnp) return;
  const int I = (int)(n % Lc.NX), J = (int)((n / Lc.NX) % Lc.NY), K = (int)(n / ((long long)Lc.NX * Lc.NY));
  double s[3] = {0.0, 0.0, 0.0};
  for (int dz = -1; dz <= 1; ++dz) {
    const int z = 2 * K + dz;
    if (z < 0 || z >= Lf.NZ) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int y = 2 * J + dy;
      if (y < 0 || y >= Lf.NY) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int x = 2 * I + dx;
        if (x < 0 || x >= Lf.NX) continue;
        const double w = w1(dx) * w1(dy) * w1(dz);
        const long long f = x + (long long)Lf.NX * (y + (long long)Lf.NY * z);
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c] += w * rf[c * Lf.Np + f];
      }
    }
  }
  const unsigned m = fmask_c[n];
#pragma unroll
  for (int c = 0; c < 3; ++c) bc[c * Lc.Np + n] = ((m >> c) & 1u) ? s[c] : 0.0;
}

// xf[fine] += sum of the (up to 8) coarse nodes around it, w * xc, free DOFs only
__global__ void k_mg_prolong_add(Lattice Lf, Lattice Lc, const double *__restrict__ xc,
                                 const uint8_t *__restrict__ fmask_f, double *__restrict__ xf) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= Lf.numnp) return;
  const int x = (int)(n % Lf.NX), y = (int)((n / Lf.NX) % Lf.NY), z = (int)(n / ((long long)Lf.NX * Lf.NY));
  const int co[3] = {x, y, z};
  const int cd[3] = {Lc.NX, Lc.NY, Lc.NZ};
  int ci[3][2], nc[3];
  double wc[3][2];
  for (int d = 0; d < 3; ++d) {
    if (co[d] & 1) {
      ci[d][0] = co[d] / 2;
      ci[d][1] = co[d] / 2 + 1;
      wc[d][0] = wc[d][1] = 0.5;
      nc[d] = ci[d][1] < cd[d] ? 2 : 1;
    } else {
      ci[d][0] = co[d] / 2;
      wc[d][0] = 1.0;
      nc[d] = 1;
    }
  }
  double s[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < nc[0]; ++a)
    for (int b = 0; b < nc[1]; ++b)
      for (int c2 = 0; c2 < nc[2]; ++c2) {
        const double w = wc[0][a] * wc[1][b] * wc[2][c2];
        const long long q = ci[0][a] + (long long)Lc.NX * (ci[1][b] + (long long)Lc.NY * ci[2][c2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c] += w * xc[c * Lc.Np + q];
      }
  const unsigned m = fmask_f[n];
#pragma unroll
  for (int c = 0; c < 3; ++c)
    if ((m >> c) & 1u) xf[c * Lf.Np + n] += s[c];
}

// coarse element = mean of its 8 children
__global__ void k_mg_young(Lattice Lf, Lattice Lc, const double *__restrict__ yf,
                           double *__restrict__ yc) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= Lc.nelx) return;
  const int I = (int)(e % Lc.nx), J = (int)((e / Lc.nx) % Lc.ny), K = (int)(e / ((long long)Lc.nx * Lc.ny));
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int x = 2 * I + (k & 1), y = 2 * J + ((k >> 1) & 1), z = 2 * K + (k >> 2);
    s += yf[x + (long long)Lf.nx * (y + (long long)Lf.ny * z)];
  }
  yc[e] = 0.125 * s;
}

// x = omega dinv b (first sweep from zero) or x += omega dinv (b - y)
__global__ void k_mg_jacobi(long long n, double omega, const double *__restrict__ dinv,
                            const double *__restrict__ b, const double *__restrict__ y,
                            double *__restrict__ x) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (y)
    x[i] += omega * dinv[i] * (b[i] - y[i]);
  else
    x[i] = omega * dinv[i] * b[i];
}

// r = b - y
__global__ void k_mg_residual(long long n, const double *__restrict__ b, const double *__restrict__ y,
                              double *__restrict__ r) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - y[i];
}

}  // namespace

void launch_mg_restrict(const Lattice &Lf, const Lattice &Lc, const double *rf,
                        const uint8_t *fmask_c, double *bc, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_restrict, dim3(blocks_for(Lc.numnp, 256)), dim3(256), 0, s, Lf, Lc, rf,
                     fmask_c, bc);
  HIPCHK(hipGetLastError());
}

void launch_mg_prolong_add(const Lattice &Lf, const Lattice &Lc, const double *xc,
                           const uint8_t *fmask_f, double *xf, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_prolong_add, dim3(blocks_for(Lf.numnp, 256)), dim3(256), 0, s, Lf, Lc,
                     xc, fmask_f, xf);
  HIPCHK(hipGetLastError());
}

void launch_mg_young(const Lattice &Lf, const Lattice &Lc, const double *yf, double *yc,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_mg_young, dim3(blocks_for(Lc.nelx, 256)), dim3(256), 0, s, Lf, Lc, yf, yc);
  HIPCHK(hipGetLastError());
}

void launch_mg_jacobi(long long n, double omega, const double *dinv, const double *b,
                      const double *y, double *x, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_jacobi, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, omega, dinv, b, y, x);
  HIPCHK(hipGetLastError());
}

void launch_mg_residual(long long n, const double *b, const double *y, double *r, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_residual, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, b, y, r);
  HIPCHK(hipGetLastError());
}

}  // namespace icarat_hip
