// k_mg.hip -- kernels of the geometric multigrid preconditioner: transfers
// between a node lattice and the lattice of the 2x coarser voxel grid, the
// coarse Young's moduli, the damped Jacobi sweep, a one-thread-per-node K.x
// with the sweep or the residual fused in, and the exact solve on the coarsest
// lattice through a dense inverse.
//
// Coarse node (I, J, K) sits on fine node (2I, 2J, 2K).  Prolongation is
// trilinear interpolation (1-D weights 1 at the coincident node, 1/2 at the
// two neighbours); restriction is its transpose.  Both are masked by the free
// DOFs of their output lattice.
#include <hip/hip_runtime.h>

#include "contract.hpp"
#include "mg_kernels.hpp"

namespace icarat_hip {

namespace {

inline unsigned blocks_for(long long n, int nt) { return (unsigned)((n + nt - 1) / nt); }

__device__ __forceinline__ double w1(int d) { return d == 0 ? 1.0 : 0.5; }

// Restriction and prolongation run fixed, unrolled stencils: indices outside
// the lattice are clamped and carry weight 0, so the loads of a plane are
// independent of each other and in flight together.
__device__ __forceinline__ void tap(int c, int dim, int d, double &w, int &q) {
  // fine coordinate c + d in [0, dim): 1-D trilinear weight (1 centre, 1/2 side)
  const int v = c + d;
  const bool in = v >= 0 && v < dim;
  w = in ? (d == 0 ? 1.0 : 0.5) : 0.0;
  q = in ? v : 0;
}

// s = sum over the 27 fine nodes around fine node (2I, 2J, 2K) of w * (rf - yf)
// (HAS_Y = false: of w * rf)
template <bool HAS_Y>
__device__ __forceinline__ void restrict_node(const Lattice &Lf, const Lattice &Lc,
                                              const double *__restrict__ rf,
                                              const double *__restrict__ yf, long long n,
                                              double (&s)[3]) {
  const int I = (int)(n % Lc.NX), J = (int)((n / Lc.NX) % Lc.NY), K = (int)(n / ((long long)Lc.NX * Lc.NY));
  double wx[3], wy[3], wz[3];
  int qx[3], qy[3], qz[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    tap(2 * I, Lf.NX, d - 1, wx[d], qx[d]);
    tap(2 * J, Lf.NY, d - 1, wy[d], qy[d]);
    tap(2 * K, Lf.NZ, d - 1, wz[d], qz[d]);
  }
  s[0] = s[1] = s[2] = 0.0;
  // one fine plane at a time: 9 nodes x 3 components (x 2 vectors) in flight
#pragma unroll 1
  for (int dz = 0; dz < 3; ++dz)
#pragma unroll
    for (int dy = 0; dy < 3; ++dy) {
      const long long row = (long long)Lf.NX * (qy[dy] + (long long)Lf.NY * qz[dz]);
      const double wyz = wy[dy] * wz[dz];
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) {
        const long long f = row + qx[dx];
        const double w = wx[dx] * wyz;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double v = HAS_Y ? rf[c * Lf.Np + f] - yf[c * Lf.Np + f] : rf[c * Lf.Np + f];
          s[c] += w * v;
        }
      }
    }
}

// (P xc) at fine node n: the 2x2x2 coarse nodes around it, weighted (a fine
// coordinate 2I coincides with coarse I: weights 1 and 0; 2I+1: 1/2 and 1/2)
__device__ __forceinline__ void prolong_node(const Lattice &Lf, const Lattice &Lc,
                                             const double *__restrict__ xc, long long n,
                                             double (&s)[3]) {
  const int co[3] = {(int)(n % Lf.NX), (int)((n / Lf.NX) % Lf.NY), (int)(n / ((long long)Lf.NX * Lf.NY))};
  const int cd[3] = {Lc.NX, Lc.NY, Lc.NZ};
  double w[3][2];
  int q[3][2];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int lo = co[d] >> 1;
    const bool odd = co[d] & 1;
    const bool hi_in = odd && lo + 1 < cd[d];
    w[d][0] = odd ? 0.5 : 1.0;
    w[d][1] = hi_in ? 0.5 : 0.0;
    q[d][0] = lo;
    q[d][1] = hi_in ? lo + 1 : lo;
  }
  s[0] = s[1] = s[2] = 0.0;
#pragma unroll
  for (int c2 = 0; c2 < 2; ++c2)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const double wt = w[0][a] * w[1][b] * w[2][c2];
        const long long qq = q[0][a] + (long long)Lc.NX * (q[1][b] + (long long)Lc.NY * q[2][c2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c] += wt * xc[c * Lc.Np + qq];
      }
}

// (K x) at node n from global memory: the 27 neighbouring displacements (zero
// outside the lattice) and the 8 surrounding elements (zero outside)
template <int SYM>
__device__ __forceinline__ void direct_node(const Lattice &L, const double *__restrict__ c_KE0,
                                            const KeVals &kv, const double *__restrict__ young,
                                            const double *__restrict__ x, long long n,
                                            double (&y)[3], double (&xc)[3]) {
  const int i = (int)(n % L.NX), j = (int)((n / L.NX) % L.NY), k = (int)(n / ((long long)L.NX * L.NY));
  double U[27][3], E[8];
#pragma unroll
  for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        const int gx = i + dx, gy = j + dy, gz = k + dz;
        const bool in = gx >= 0 && gx < L.NX && gy >= 0 && gy < L.NY && gz >= 0 && gz < L.NZ;
        const long long m = in ? gx + (long long)L.NX * (gy + (long long)L.NY * gz) : 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) U[nb(dx, dy, dz)][c] = in ? x[c * L.Np + m] : 0.0;
      }
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int ex = i + (s & 1) - 1, ey = j + ((s >> 1) & 1) - 1, ez = k + ((s >> 2) & 1) - 1;
    const bool in = ex >= 0 && ex < L.nx && ey >= 0 && ey < L.ny && ez >= 0 && ez < L.nz;
    E[s] = in ? young[ex + (long long)L.nx * (ey + (long long)L.ny * ez)] : 0.0;
  }
  node_contract<SYM>(c_KE0, kv, U, E, y);
#pragma unroll
  for (int c = 0; c < 3; ++c) xc[c] = U[13][c];
}

// out[n] from (K x)[n] under a direct mode, masked
template <int MODE>
__device__ __forceinline__ void direct_out(long long Np, long long n, unsigned m, const double (&y)[3],
                                           const double (&xc)[3], const double *__restrict__ b,
                                           const double *__restrict__ dinv, double omega,
                                           double *__restrict__ out) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const long long q = c * Np + n;
    double v;
    if constexpr (MODE == MG_APPLY) v = y[c];
    else if constexpr (MODE == MG_JACOBI) v = xc[c] + omega * dinv[q] * (b[q] - y[c]);
    else v = b[q] - y[c];
    out[q] = ((m >> c) & 1u) ? v : 0.0;
  }
}

template <bool HAS_Y>
__global__ void k_mg_restrict(Lattice Lf, Lattice Lc, const double *__restrict__ rf,
                              const double *__restrict__ yf, const uint8_t *__restrict__ fmask_c,
                              double *__restrict__ bc, const double *__restrict__ dinv_c,
                              double omega, double *__restrict__ xc) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= Lc.numnp) return;
  double s[3];
  restrict_node<HAS_Y>(Lf, Lc, rf, yf, n, s);
  const unsigned m = fmask_c[n];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double v = ((m >> c) & 1u) ? s[c] : 0.0;
    bc[c * Lc.Np + n] = v;
    if (dinv_c) xc[c * Lc.Np + n] = omega * dinv_c[c * Lc.Np + n] * v;
  }
}

__global__ void k_mg_prolong_add(Lattice Lf, Lattice Lc, const double *__restrict__ xc,
                                 const uint8_t *__restrict__ fmask_f, double *__restrict__ xf) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= Lf.numnp) return;
  double s[3];
  prolong_node(Lf, Lc, xc, n, s);
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

constexpr int DOT_BLOCKS = 2048;

__global__ __launch_bounds__(256) void k_mg_jacobi_dot(long long n, double omega,
                                                       const double *__restrict__ dinv,
                                                       const double *__restrict__ b,
                                                       const double *__restrict__ y,
                                                       double *__restrict__ x,
                                                       const double *__restrict__ q,
                                                       double *part) {
  __shared__ double sh[8];
  double zr = 0.0, zq = 0.0;
  const long long stride = (long long)gridDim.x * 256;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += stride) {
    const double bi = b[i];
    const double xi = x[i] + omega * dinv[i] * (bi - y[i]);
    x[i] = xi;
    zr += xi * bi;
    zq += xi * q[i];
  }
  double v[2] = {zr, zq};
  store_partials<256, 2>(v, part, sh);
}

// r = b - y
__global__ void k_mg_residual(long long n, const double *__restrict__ b, const double *__restrict__ y,
                              double *__restrict__ r) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - y[i];
}

template <int SYM, int MODE>
__global__ __launch_bounds__(256) void k_mg_direct(Lattice L, const double *__restrict__ c_KE0,
                                                   KeVals kv, const double *__restrict__ young,
                                                   const uint8_t *__restrict__ fmask,
                                                   const double *__restrict__ x,
                                                   const double *__restrict__ b,
                                                   const double *__restrict__ dinv, double omega,
                                                   double *__restrict__ out) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= L.numnp) return;
  double y[3], xc[3];
  direct_node<SYM>(L, c_KE0, kv, young, x, n, y, xc);
  direct_out<MODE>(L.Np, n, fmask[n], y, xc, b, dinv, omega, out);
}

// ---- dense exact solve on the coarsest lattice -----------------------------

__global__ void k_mg_dense_zero(long long nn, double *__restrict__ A) {
  const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (q < nn) A[q] = 0.0;
}

// Scatter of the elements of one parity class (ex, ey, ez mod 2 = px, py, pz):
// they share no node, so no two threads add to the same entry.
__global__ void k_mg_dense_scatter(Lattice L, const double *__restrict__ ke,
                                   const double *__restrict__ young, int px, int py, int pz,
                                   double *__restrict__ A) {
  const int mx = (L.nx - px + 1) / 2, my = (L.ny - py + 1) / 2, mz = (L.nz - pz + 1) / 2;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= (long long)mx * my * mz * 576) return;
  const long long n = 3 * L.numnp;
  const int ent = (int)(t % 576);
  const long long m = t / 576;
  const int ex = px + 2 * (int)(m % mx), ey = py + 2 * (int)((m / mx) % my), ez = pz + 2 * (int)(m / ((long long)mx * my));
  const int r = ent / 24, c = ent % 24;
  const int a = r / 3, i = r % 3, b = c / 3, j = c % 3;
  const long long na = (ex + lnode_x(a)) + (long long)L.NX * ((ey + lnode_y(a)) + (long long)L.NY * (ez + lnode_z(a)));
  const long long nb_ = (ex + lnode_x(b)) + (long long)L.NX * ((ey + lnode_y(b)) + (long long)L.NY * (ez + lnode_z(b)));
  const double E = young[ex + (long long)L.nx * (ey + (long long)L.ny * ez)];
  A[(i * L.numnp + na) * n + j * L.numnp + nb_] += E * ke[ent];
}

// constrained DOFs: identity rows and columns
__global__ void k_mg_dense_constrain(Lattice L, const uint8_t *__restrict__ fmask,
                                     double *__restrict__ A) {
  const long long n = 3 * L.numnp;
  const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (q >= n * n) return;
  const long long r = q / n, c = q % n;
  const bool fr = (fmask[r % L.numnp] >> (r / L.numnp)) & 1u;
  const bool fc = (fmask[c % L.numnp] >> (c / L.numnp)) & 1u;
  if (!fr || !fc) A[q] = r == c ? 1.0 : 0.0;
}

// column-major: A(i, j) = A[i + j n]; the lower triangle (i > j) is valid
__global__ void k_mg_dense_symmetrize(int n, double *__restrict__ A) {
  const long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (q >= (long long)n * n) return;
  const int i = (int)(q % n), j = (int)(q / n);
  if (i < j) A[q] = A[j + (long long)i * n];
}

// x = A b: one block per row, threads over the columns, block reduction
__global__ __launch_bounds__(256) void k_mg_dense_solve(Lattice L, const double *__restrict__ A,
                                                        const double *__restrict__ b,
                                                        double *__restrict__ x) {
  __shared__ double sh[4];
  const int n = (int)(3 * L.numnp), np1 = (int)L.numnp;
  const int row = blockIdx.x;
  double s = 0.0;
  for (int j = threadIdx.x; j < n; j += 256)
    s += A[(long long)row * n + j] * b[(long long)(j / np1) * L.Np + j % np1];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) x[(long long)(row / np1) * L.Np + row % np1] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

template <int SYM>
void direct_dispatch(int mode, const Lattice &L, const KeCoef &kc, const double *young,
                     const uint8_t *fmask, const double *x, const double *b, const double *dinv,
                     double omega, double *out, hipStream_t s) {
  const dim3 g(blocks_for(L.numnp, 256)), t(256);
  if (mode == MG_APPLY)
    hipLaunchKernelGGL((k_mg_direct<SYM, MG_APPLY>), g, t, 0, s, L, kc.ke, kc.vals, young, fmask, x,
                       b, dinv, omega, out);
  else if (mode == MG_JACOBI)
    hipLaunchKernelGGL((k_mg_direct<SYM, MG_JACOBI>), g, t, 0, s, L, kc.ke, kc.vals, young, fmask,
                       x, b, dinv, omega, out);
  else
    hipLaunchKernelGGL((k_mg_direct<SYM, MG_RESIDUAL>), g, t, 0, s, L, kc.ke, kc.vals, young,
                       fmask, x, b, dinv, omega, out);
}

}  // namespace

void launch_mg_restrict(const Lattice &Lf, const Lattice &Lc, const double *rf, const double *yf,
                        const uint8_t *fmask_c, double *bc, const double *dinv_c, double omega,
                        double *xc, hipStream_t s) {
  if (yf)
    hipLaunchKernelGGL(k_mg_restrict<true>, dim3(blocks_for(Lc.numnp, 256)), dim3(256), 0, s, Lf, Lc,
                       rf, yf, fmask_c, bc, dinv_c, omega, xc);
  else
    hipLaunchKernelGGL(k_mg_restrict<false>, dim3(blocks_for(Lc.numnp, 256)), dim3(256), 0, s, Lf, Lc,
                       rf, yf, fmask_c, bc, dinv_c, omega, xc);
  HIPCHK(hipGetLastError());
}

void launch_mg_dense_assemble(const Lattice &L, const double *ke, const double *young,
                              const uint8_t *fmask, double *A, hipStream_t s) {
  const long long n = 3 * L.numnp;
  if (n > MG_DENSE_MAX) throw Error(kInvalidArgument, "lattice too large for the dense solve");
  hipLaunchKernelGGL(k_mg_dense_zero, dim3(blocks_for(n * n, 256)), dim3(256), 0, s, n * n, A);
  for (int cls = 0; cls < 8; ++cls) {
    const int px = cls & 1, py = (cls >> 1) & 1, pz = cls >> 2;
    const long long cnt = (long long)((L.nx - px + 1) / 2) * ((L.ny - py + 1) / 2) * ((L.nz - pz + 1) / 2) * 576;
    if (cnt == 0) continue;
    hipLaunchKernelGGL(k_mg_dense_scatter, dim3(blocks_for(cnt, 256)), dim3(256), 0, s, L, ke, young,
                       px, py, pz, A);
  }
  hipLaunchKernelGGL(k_mg_dense_constrain, dim3(blocks_for(n * n, 256)), dim3(256), 0, s, L, fmask, A);
  HIPCHK(hipGetLastError());
}

void launch_mg_dense_symmetrize(int n, double *A, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_dense_symmetrize, dim3(blocks_for((long long)n * n, 256)), dim3(256), 0, s,
                     n, A);
  HIPCHK(hipGetLastError());
}

void launch_mg_dense_solve(const Lattice &L, const double *ainv, const double *b, double *x,
                           hipStream_t s) {
  if (3 * L.numnp > MG_DENSE_MAX) throw Error(kInvalidArgument, "lattice too large for the dense solve");
  hipLaunchKernelGGL(k_mg_dense_solve, dim3((unsigned)(3 * L.numnp)), dim3(256), 0, s, L, ainv, b, x);
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

unsigned launch_mg_jacobi_dot(long long n, double omega, const double *dinv, const double *b,
                              const double *y, double *x, const double *q, double *part,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_mg_jacobi_dot, dim3(DOT_BLOCKS), dim3(256), 0, s, n, omega, dinv, b, y, x, q,
                     part);
  HIPCHK(hipGetLastError());
  return DOT_BLOCKS;
}

void launch_mg_residual(long long n, const double *b, const double *y, double *r, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_residual, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, b, y, r);
  HIPCHK(hipGetLastError());
}

void launch_mg_direct(int mode, const Lattice &L, const KeCoef &kc, const double *young,
                      const uint8_t *fmask, const double *x, const double *b, const double *dinv,
                      double omega, double *out, hipStream_t s) {
  if (mode < MG_APPLY || mode > MG_RESIDUAL) throw Error(kInvalidArgument, "unknown direct mode");
  if (kc.sym == 2)
    direct_dispatch<2>(mode, L, kc, young, fmask, x, b, dinv, omega, out, s);
  else if (kc.sym == 1)
    direct_dispatch<1>(mode, L, kc, young, fmask, x, b, dinv, omega, out, s);
  else
    direct_dispatch<0>(mode, L, kc, young, fmask, x, b, dinv, omega, out, s);
  HIPCHK(hipGetLastError());
}

}  // namespace icarat_hip
